// Dispatch and launch of the element-operator kernels (the kernel template
// is in nb_ax_kernel.h; its instantiations in nb_ax_i*.hip).
#include "nb_ax_kernel.h"

namespace nb {
namespace {

using axk::AxArgs;
using axk::AxFn;
using axk::Key;
using axk::Table;

Table &table()
{
    static Table t;
    static bool done = false;
    if (!done) {
        done = true;
        axk::reg_low(t);
        axk::reg_high(t);
        axk::reg_6(t);
        axk::reg_8(t);
        axk::reg_10(t);
        axk::reg_12(t);
    }
    return t;
}

AxFn lookup(int nx, int epb, bool fma, int mode, int pf, int xu)
{
    auto it = table().find(Key(nx, epb, fma ? 1 : 0, mode, pf, mode == 0 ? 0 : xu));
    return it == table().end() ? nullptr : it->second;
}

constexpr int round64(int x) { return (x + 63) / 64 * 64; }

}  // namespace

AxVariant ax_default_variant(int nx1)
{
    AxVariant v;
    // measured fastest at N=9 on MI355X: one element per workgroup at this
    // order, factors loaded one t-step ahead of their use where that shape
    // is compiled (else in the geometric phase), x update fused
    v.epb = axk::epb_for(nx1);
    v.pf = lookup(nx1, v.epb, false, 1, 3, 1) ? 3 : 0;
    v.xu = 1;
    return v;
}

std::vector<AxVariant> ax_variants(int nx1)
{
    std::vector<AxVariant> out;
    for (const auto &kv : table()) {
        int nx, epb, fma, mode, pf, xu;
        std::tie(nx, epb, fma, mode, pf, xu) = kv.first;
        if (nx == nx1 && fma == 0 && mode == 1) out.push_back(AxVariant{epb, pf, xu});
    }
    return out;
}

int ax_blocks(const nekbone_problem *P)
{
    const int epb = P->axv.epb;
    return (P->nelt + epb - 1) / epb;
}

int launch_ax_local(nekbone_problem *P, double *w, const double *u)
{
    const bool fma = P->arith == NEKBONE_ARITH_FMA;
    int epb = P->axv.epb, pf = P->axv.pf;
    AxFn fn = lookup(P->nx1, epb, fma, 0, pf, 0);
    if (!fn) {
        epb = axk::epb_for(P->nx1);
        pf = 0;
        fn = lookup(P->nx1, epb, fma, 0, pf, 0);
    }
    if (!fn) return fail(NEKBONE_ERR_UNSUPPORTED, "ax: nx1 outside 2..16");
    AxArgs A{};
    A.w = w;
    A.u = u;
    A.g = P->d_g;
    A.gcs = P->n;
    A.Drow = P->d_Drow;
    A.Dcol = P->d_D;
    A.iref = P->d_iref;
    A.nelt = P->nelt;
    const int threads = round64(P->nx1 * P->nx1 * epb);
    const int blocks = (P->nelt + epb - 1) / epb;
    if (int rc = timed_begin(P, 0)) return rc;
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(threads), 0, stream(), A);
    NB_HIP(hipGetLastError());
    return timed_end(P, 0);
}

int launch_ax_cg(nekbone_problem *P, double *x, const double *z, int mbit, int rev, int gx,
                 double *pin, double *pout)
{
    // x-ring (pout set): the variant without the x update
    AxFn fn = lookup(P->nx1, P->axv.epb, P->arith == NEKBONE_ARITH_FMA, 1, P->axv.pf,
                     pout ? 0 : P->axv.xu);
    if (!fn) return fail(NEKBONE_ERR_UNSUPPORTED, "ax: no kernel for this order/variant");
    AxArgs A{};
    A.w = P->d_w;
    A.p = pout ? pin : P->d_p;
    A.pout = pout;
    A.r = z;
    A.x = x;
    A.g = P->d_g;
    A.gcs = P->n;
    A.Drow = P->d_Drow;
    A.Dcol = P->d_D;
    A.flag = P->d_flag;
    A.iref = P->d_iref;
    A.mbit = mbit;
    A.scal = P->d_scal;
    A.part = P->d_part;
    A.nelt = P->nelt;
    A.rev = rev;
    A.gx = gx;
    const int threads = round64(P->nx1 * P->nx1 * P->axv.epb);
    if (int rc = timed_begin(P, 0)) return rc;
    hipLaunchKernelGGL(fn, dim3(ax_blocks(P)), dim3(threads), 0, stream(), A);
    NB_HIP(hipGetLastError());
    return timed_end(P, 0);
}

int launch_ax_flush(nekbone_problem *P, float *x32, const float *p32, int gate)
{
    const int epb = axk::epb_for(P->nx1);
    AxFn fn = lookup(P->nx1, epb, false, 3, 3, 0);
    if (!fn) fn = lookup(P->nx1, epb, false, 3, 0, 0);
    if (!fn || P->arith == NEKBONE_ARITH_FMA) return NEKBONE_ERR_UNSUPPORTED;
    AxArgs A{};
    A.w = P->d_w;
    A.x = P->d_x;
    A.g = P->d_g;
    A.gcs = P->n;
    A.Drow = P->d_Drow;
    A.Dcol = P->d_D;
    A.iref = P->d_iref;
    A.scal = P->d_scal;
    A.nelt = P->nelt;
    A.x32 = x32;
    A.p32 = p32;
    A.gate = gate;
    const int threads = round64(P->nx1 * P->nx1 * epb);
    const int blocks = (P->nelt + epb - 1) / epb;
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(threads), 0, stream(), A);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

}  // namespace nb
