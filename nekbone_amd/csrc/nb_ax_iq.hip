// Work-queue element kernels of the batched CG (k_ax_queue): the tile sets
// the batch autotuner chooses among.  Elements per workgroup at orders 6,
// 8, 10, 12 are shapes the single-problem kernel table also has, so a
// problem solved alone with the same shape gets the same values.
#include "nb_ax_kernel.h"

namespace nb {
namespace axk {
namespace {
template <int NTB, int E6, int E8, int E10, int E12>
QShape shape()
{
    QShape q{NTB, {E6, E8, E10, E12}, {{nullptr, nullptr}, {nullptr, nullptr}}};
    q.fn[0][0] = k_ax_queue<NTB, false, false, E6, E8, E10, E12>;
    q.fn[0][1] = k_ax_queue<NTB, false, true, E6, E8, E10, E12>;
    q.fn[1][0] = k_ax_queue<NTB, true, false, E6, E8, E10, E12>;
    q.fn[1][1] = k_ax_queue<NTB, true, true, E6, E8, E10, E12>;
    return q;
}
}  // namespace

void regq_all(std::vector<QShape> &t)
{
    t.push_back(shape<192, 2, 2, 1, 1>());  // the orders' default shapes
    t.push_back(shape<192, 4, 2, 1, 1>());
    t.push_back(shape<256, 4, 4, 2, 1>());
    t.push_back(shape<320, 4, 4, 2, 2>());  // every order doubled
}
}  // namespace axk
}  // namespace nb
