// Element-operator kernels for nx1 = 2..5, 7, 9 (default shape).
#include "nb_ax_kernel.h"

namespace nb {
namespace axk {
void reg_low(Table &t)
{
    reg_order<2>(t, false);
    reg_order<3>(t, false);
    reg_order<4>(t, false);
    reg_order<5>(t, false);
    reg_order<7>(t, false);
    reg_order<9>(t, false);
}
}  // namespace axk
}  // namespace nb
