// Element-operator kernels for nx1 = 6 (all shapes the autotuner may pick).
#include "nb_ax_kernel.h"

namespace nb {
namespace axk {
void reg_6(Table &t) { reg_order<6>(t, true); }
}  // namespace axk
}  // namespace nb
