// FP16 matrix-core element kernels of the mixed-precision solver, nx1 = 4..16.
#include "nb_ax_kernel.h"

namespace nb {
namespace axk {
void reg32_mf_all(Table32 &t)
{
    reg32_mf<4>(t);
    reg32_mf<5>(t);
    reg32_mf<6>(t);
    reg32_mf<7>(t);
    reg32_mf<8>(t);
    reg32_mf<9>(t);
    reg32_mf<10>(t);
    reg32_mf<11>(t);
    reg32_mf<12>(t);
    reg32_mf<13>(t);
    reg32_mf<14>(t);
    reg32_mf<15>(t);
    reg32_mf<16>(t);
}
}  // namespace axk
}  // namespace nb
