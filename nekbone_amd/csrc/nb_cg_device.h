// Device helpers of the CG kernels: fixed-order reductions of partial
// sums and the scalar bookkeeping of an iteration (alpha, beta, rtr,
// convergence), shared by the single-problem and the batched kernels and
// by the mixed-precision solver.
#pragma once

#include "nb_device.h"

namespace nb {

constexpr int kFinalThreads = 1024;

// Template value of PRE for the fused Jacobi CG in z form: the vector the
// iterations carry in place of r is z = M^-1 r (updated as z -= alpha
// M^-1 w), so the element kernel reads z alone; r = z / dinv where the dot
// products need it and once at the end.
constexpr int kPreJacobiZ = 3;
constexpr bool pre_jacobi(int pre) { return pre == NEKBONE_PRECOND_JACOBI || pre == kPreJacobiZ; }

// c = 1/multiplicity (exactly the reference's 1./c(i)); the common hex
// multiplicities are powers of two, exact without a division.
__device__ __forceinline__ double inv_mult(uint8_t fl)
{
    const int m = fl >> kMultShift;
    return m == 1 ? 1.0 : m == 2 ? 0.5 : m == 4 ? 0.25 : m == 8 ? 0.125 : 1. / double(m);
}

// Fixed-order sum of a strided slice with sixteen interleaved accumulators
// (sixteen loads in flight per thread: the single-block final kernels sum
// up to ~60k partials, and their time is the memory latency of the
// rounds of loads), combined pairwise in a fixed tree.
template <int NT>
__device__ __forceinline__ double slice_sum(const double *__restrict__ a, int64_t na)
{
    constexpr int U = 16;
    double s[U];
#pragma unroll
    for (int u = 0; u < U; ++u) s[u] = 0.0;
    int64_t i = threadIdx.x;
    for (; i + (U - 1) * NT < na; i += U * NT) {
#pragma unroll
        for (int u = 0; u < U; ++u) s[u] = s[u] + a[i + u * NT];
    }
    for (int u = 0; i < na; i += NT, ++u) s[u] = s[u] + a[i];
#pragma unroll
    for (int h = U / 2; h > 0; h >>= 1)
#pragma unroll
        for (int u = 0; u < h; ++u) s[u] = s[u] + s[u + h];
    return s[0];
}

// Sum of a partials region in a fixed order by one block.
template <int NT>
__device__ __forceinline__ void sum_region2(const double *__restrict__ a, int64_t na,
                                            const double *__restrict__ b, int64_t nb, double &s0,
                                            double &s1)
{
    double a0 = slice_sum<NT>(a, na);
    double a1 = b ? slice_sum<NT>(b, nb) : 0.0;
    block_sum2<NT>(a0, a1);
    s0 = a0;
    s1 = a1;
}

__device__ __forceinline__ void prepare_next(CgScalars *s, double *hist, int hcap)
{
    s->rtz2 = s->rtz1;
    s->rtz1 = s->rtz_next;
    s->beta = (s->iter == 0) ? 0.0 : s->rtz1 / s->rtz2;
    const int h = 1 + 3 * s->iter;
    if (h < hcap) hist[h] = s->rtz1;
}

template <int NT>
__device__ __forceinline__ void pap_final_body(const double *__restrict__ part, int64_t npart,
                                               CgScalars *__restrict__ s,
                                               double *__restrict__ hist, int hcap)
{
    if (s->done) return;
    double pap, dummy;
    sum_region2<NT>(part, npart, nullptr, 0, pap, dummy);
    if (threadIdx.x == 0) {
        s->pap = pap;
        s->alpha = s->rtz1 / pap;
        s->xpending = 0;  // the element kernel of this iteration applied it
        if (!(pap > 0.0) || !isfinite(pap)) {
            s->status = 1;
            if (s->tol2 > 0.0) s->done = 1;
        }
        const int h = 2 + 3 * s->iter;
        if (h < hcap) hist[h] = pap;
    }
}

template <int PRE, int NT>
__device__ __forceinline__ void update_final_body(const double *__restrict__ part, int nblk,
                                                  CgScalars *__restrict__ s,
                                                  double *__restrict__ hist, int hcap,
                                                  int set_xpending)
{
    if (s->done) return;
    double rtr, rtz;
    sum_region2<NT>(part, nblk, part + nblk, nblk, rtr, rtz);
    if (threadIdx.x == 0) {
        s->rtr = rtr;
        const int h = 3 + 3 * s->iter;
        if (h < hcap) hist[h] = rtr;
        s->iter += 1;
        if (set_xpending) s->xpending = 1;
        s->rtz_next = pre_jacobi(PRE) ? rtz : rtr;
        if (s->tol2 > 0.0 && rtr <= s->tol2 * s->rtr0) s->done = 1;
        prepare_next(s, hist, hcap);
    }
}

}  // namespace nb
