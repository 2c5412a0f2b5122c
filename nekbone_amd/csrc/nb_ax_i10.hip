// Element-operator kernels for nx1 = 10 (all shapes the autotuner may pick).
#include "nb_ax_kernel.h"

namespace nb {
namespace axk {
void reg_10(Table &t) { reg_order<10>(t, true); }
}  // namespace axk
}  // namespace nb
