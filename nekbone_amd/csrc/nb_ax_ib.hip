// Batched element-operator kernels for the CG modes, every order 2..16,
// default shape (elements per workgroup of the order, factors loaded in
// the geometric phase, with and without the deferred x update).
#include "nb_ax_kernel.h"

namespace nb {
namespace axk {
void regb_all(TableB &t)
{
    regb_order<2>(t);
    regb_order<3>(t);
    regb_order<4>(t);
    regb_order<5>(t);
    regb_order<6>(t);
    regb_order<7>(t);
    regb_order<8>(t);
    regb_order<9>(t);
    regb_order<10>(t);
    regb_order<11>(t);
    regb_order<12>(t);
    regb_order<13>(t);
    regb_order<14>(t);
    regb_order<15>(t);
    regb_order<16>(t);
}
}  // namespace axk
}  // namespace nb
