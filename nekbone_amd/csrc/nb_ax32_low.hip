// FP32 element-operator kernels for nx1 = 2..9 (mixed-precision inner
// solve); every shape for nx1 = 8.
#include "nb_ax_kernel.h"

namespace nb {
namespace axk {
void reg32_low(Table32 &t)
{
    reg32_order<2>(t, false);
    reg32_order<3>(t, false);
    reg32_order<4>(t, false);
    reg32_order<5>(t, false);
    reg32_order<6>(t, false);
    reg32_order<7>(t, false);
    reg32_order<8>(t, true);
    reg32_order<9>(t, false);
}
}  // namespace axk
}  // namespace nb
