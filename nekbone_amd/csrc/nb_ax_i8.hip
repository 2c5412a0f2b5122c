// Element-operator kernels for nx1 = 8 (all shapes the autotuner may pick).
#include "nb_ax_kernel.h"

namespace nb {
namespace axk {
void reg_8(Table &t) { reg_order<8>(t, true); }
}  // namespace axk
}  // namespace nb
