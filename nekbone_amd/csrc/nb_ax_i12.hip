// Element-operator kernels for nx1 = 12 (all shapes the autotuner may pick).
#include "nb_ax_kernel.h"

namespace nb {
namespace axk {
void reg_12(Table &t) { reg_order<12>(t, true); }
}  // namespace axk
}  // namespace nb
