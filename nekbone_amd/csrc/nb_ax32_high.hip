// FP32 element-operator kernels for nx1 = 10..16 (mixed-precision inner
// solve); every shape for nx1 = 10.
#include "nb_ax_kernel.h"

namespace nb {
namespace axk {
void reg32_high(Table32 &t)
{
    reg32_order<10>(t, true);
    reg32_order<11>(t, false);
    reg32_order<12>(t, false);
    reg32_order<13>(t, false);
    reg32_order<14>(t, false);
    reg32_order<15>(t, false);
    reg32_order<16>(t, false);
}
}  // namespace axk
}  // namespace nb
