// Element-operator kernels for nx1 = 11, 13..16 (default shape).
#include "nb_ax_kernel.h"

namespace nb {
namespace axk {
void reg_high(Table &t)
{
    reg_order<11>(t, false);
    reg_order<13>(t, false);
    reg_order<14>(t, false);
    reg_order<15>(t, false);
    reg_order<16>(t, false);
}
}  // namespace axk
}  // namespace nb
