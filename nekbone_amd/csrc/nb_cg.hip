// CG (src/cg.f:2-84) as device kernels with all scalars on the device.
//
// Fused path (c == 1/multiplicity, which is what set_multiplicity,
// src/driver.f:274-291, produces), per iteration:
//   1. k_ax<...,MODE>   x += alpha_prev p_prev; p = beta p + z; w_L = A_L p;
//                       unshared points finished (+0.1p, mask, pap part)
//   2. k_gs_cls<true>   shared groups: sum in ascending local index
//                       (src/jl/gs_local.c:60-87), then +0.1p, mask and
//                       pap part per copy
//   3. k_pap_final      pap, alpha = rtz1/pap
//   4. k_update         r += (-alpha) w; parts of rtr = (r,c,r) and of
//                       rtz = (r,c,z) for the next iteration
//   5. k_update_final   rtr, beta for the next iteration, convergence
// Values are bit-identical to the reference's sequence; only the
// summation order of the dot products differs (fixed tree, so results are
// reproducible run to run).
//
// Generic path (any c): the reference's sequence with separate kernels
// (p update, ax, dssum, shift/mask/pap, x/r update with rtr).
#include "nb_cg_device.h"

namespace nb {
namespace {

// x = 0, r = mask(f) (src/cg.f:38-40); parts of glsc3(r,c,r) and of
// glsc3(r,c,z) with z = M^-1 r.  PRE = kPreJacobiZ stores z in place of r.
template <int PRE>
__global__ void __launch_bounds__(kRedThreads)
    k_cg_init(double *__restrict__ x, double *__restrict__ r, const double *__restrict__ f,
              const double *__restrict__ c, const double *__restrict__ dinv,
              const uint8_t *__restrict__ flag, int mbit, int64_t n, double *__restrict__ part,
              int nblk)
{
    double a0 = 0.0, a1 = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kRedThreads;
    for (int64_t i = int64_t(blockIdx.x) * kRedThreads + threadIdx.x; i < n; i += stride) {
        x[i] = 0.0;
        const double v = (flag[i] & mbit) ? 0.0 : f[i];
        const double rc = v * c[i];
        a0 = a0 + rc * v;
        if constexpr (pre_jacobi(PRE)) {
            const double z = v * dinv[i];
            a1 = a1 + rc * z;
            r[i] = PRE == kPreJacobiZ ? z : v;
        } else {
            r[i] = v;
        }
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a0;
        part[nblk + blockIdx.x] = a1;
    }
}

template <int PRE>
__global__ void __launch_bounds__(kFinalThreads)
    k_cg_init_final(const double *__restrict__ part, int nblk, CgScalars *__restrict__ s,
                    double *__restrict__ hist, int hcap, double tol2)
{
    double rtr, rtz;
    sum_region2<kFinalThreads>(part, nblk, part + nblk, nblk, rtr, rtz);
    if (threadIdx.x == 0) {
        s->rtr0 = rtr;
        s->rtr = rtr;
        s->rtz1 = 1.0;  // src/cg.f:36
        s->rtz_next = pre_jacobi(PRE) ? rtz : rtr;
        s->iter = 0;
        s->alpha = 0.0;
        s->done = (tol2 > 0.0 && rtr == 0.0) ? 1 : 0;
        s->status = 0;
        s->xpending = 0;
        s->xdone = 0;
        s->tol2 = tol2;
        if (hcap > 0) hist[0] = rtr;
        prepare_next(s, hist, hcap);
    }
}

__global__ void __launch_bounds__(kFinalThreads)
    k_pap_final(const double *__restrict__ part, int64_t npart, CgScalars *__restrict__ s,
                double *__restrict__ hist, int hcap)
{
    pap_final_body<kFinalThreads>(part, npart, s, hist, hcap);
}

// [XP] x += alpha p (add2s2, src/cg.f:63) when the element kernel does not
// carry it; r += (-alpha) w (add2s2, src/cg.f:64); parts of rtr = (r*c)*r
// (src/cg.f:66) and of the next rtz = (r*c)*z.  Logical block blk owns
// one contiguous chunk of the vectors, so the kernel walks memory in
// block order (either direction, see sweep_block) and each partial sums
// the same points in the same order whatever the direction.  Eight
// independent points per thread per step keep enough loads in flight
// (measured 1% faster than four at 32^3 elements, N=9).
template <int PRE, bool FMA, bool XP>
__device__ __forceinline__ void update_body(double *__restrict__ r, const double *__restrict__ w,
                                            const double *__restrict__ dinv,
                                            const uint8_t *__restrict__ flag,
                                            double *__restrict__ x, const double *__restrict__ p,
                                            int64_t n, double *__restrict__ part, int nblk,
                                            const CgScalars *__restrict__ s, const int blk)
{
    if (s->done) return;
    const double alpha = s->alpha, alphm = -alpha;
    double a0 = 0.0, a1 = 0.0;
    constexpr int U = 8;
    constexpr int64_t step = int64_t(kRedThreads) * U;
    const int64_t cs = (n + int64_t(nblk) * step - 1) / (int64_t(nblk) * step) * step;
    const int64_t beg = int64_t(blk) * cs, end = beg + cs < n ? beg + cs : n;
    for (int64_t i0 = beg + threadIdx.x; i0 < end; i0 += step) {
        double rv[U], wv[U], dv[U], xv[U], pv[U];
        uint8_t fv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + int64_t(u) * kRedThreads;
            if (i < end) {
                rv[u] = r[i];
                wv[u] = w[i];
                fv[u] = flag[i];
                if constexpr (pre_jacobi(PRE)) dv[u] = dinv[i];
                if constexpr (XP) {
                    xv[u] = x[i];
                    pv[u] = p[i];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + int64_t(u) * kRedThreads;
            if (i < end) {
                if constexpr (XP) {
                    if constexpr (FMA)
                        x[i] = __builtin_fma(alpha, pv[u], xv[u]);
                    else
                        x[i] = xv[u] + alpha * pv[u];
                }
                if constexpr (PRE == kPreJacobiZ) {
                    // z form: rv holds z; z -= alpha dinv w, r = z / dinv
                    double zi;
                    if constexpr (FMA)
                        zi = __builtin_fma(alphm, dv[u] * wv[u], rv[u]);
                    else
                        zi = rv[u] + alphm * (dv[u] * wv[u]);
                    r[i] = zi;
                    const double ri = zi / dv[u];
                    const double rc = ri * inv_mult(fv[u]);
                    a0 = a0 + rc * ri;
                    a1 = a1 + rc * zi;
                } else {
                    double ri;
                    if constexpr (FMA)
                        ri = __builtin_fma(alphm, wv[u], rv[u]);
                    else
                        ri = rv[u] + alphm * wv[u];
                    r[i] = ri;
                    const double rc = ri * inv_mult(fv[u]);
                    a0 = a0 + rc * ri;
                    if constexpr (PRE == NEKBONE_PRECOND_JACOBI) a1 = a1 + rc * (ri * dv[u]);
                }
            }
        }
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[blk] = a0;
        part[nblk + blk] = a1;
    }
}

template <int PRE, bool FMA, bool XP>
__global__ void __launch_bounds__(kRedThreads)
    k_update(double *__restrict__ r, const double *__restrict__ w,
             const double *__restrict__ dinv, const uint8_t *__restrict__ flag,
             double *__restrict__ x, const double *__restrict__ p, int64_t n,
             double *__restrict__ part, int nblk, const CgScalars *__restrict__ s, int rev)
{
    update_body<PRE, FMA, XP>(r, w, dinv, flag, x, p, n, part, nblk, s,
                              sweep_block(blockIdx.x, nblk, rev));
}

// Gather-form update, one workgroup per G.eb consecutive elements (their
// table rows staged in LDS): for shared points w is formed here as the
// sum of the copies in member order + 0.1 p, masked (what the
// gather-scatter kernel stores), then the update of update_body.  The
// member loads depend only on the point's position and the LDS table, so
// they are issued with the points' own loads.
//
// [XF] x-ring flush (iterations k with k % kRing == kRing - 1): x receives
// alpha_j p_j of the iterations j in [xdone, k] from the ring (rs apart),
// in iteration order.
template <int PRE, bool FMA, bool XP, bool XF>
__device__ __forceinline__ void update_gx_body(double *__restrict__ r, const double *__restrict__ w,
                                               const double *__restrict__ dinv,
                                               const uint8_t *__restrict__ flag,
                                               double *__restrict__ x,
                                               const double *__restrict__ p, int nelt,
                                               double *__restrict__ part, int nblk,
                                               const CgScalars *__restrict__ s, const int blk,
                                               const GxView &G, const double *__restrict__ ring,
                                               int64_t rs)
{
    __shared__ int32_t tabs[kGxMaxEb * kGxStride];
    if (s->done) return;
    const int64_t e0 = int64_t(blk) * G.eb;
    const int ne = int(nelt - e0 < G.eb ? nelt - e0 : G.eb);
    gx_rows<kRedThreads>(tabs, G, e0, ne);
    __shared__ double vsum[kGxMaxEb * 8];
    gx_vertices<double, kRedThreads>(vsum, w, tabs, ne);
    const int np = ne * G.n3;
    const int64_t beg = e0 * G.n3;
    const double alpha = s->alpha, alphm = -alpha;
    double a0 = 0.0, a1 = 0.0;
    constexpr int U = kGxU;
    for (int j0 = threadIdx.x; j0 < np; j0 += U * kRedThreads) {
        double rv[U], wv[U], dv[U], xv[U], pv[U], ps[U];
        uint8_t fv[U];
        bool sh[U], gm[U];
        double gv[U][4];
        int gc[U], go[U], gs[U], ge[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * kRedThreads;
            sh[u] = false;
            if (j < np) {
                const int64_t i = beg + j;
                rv[u] = r[i];
                wv[u] = w[i];
                fv[u] = flag[i];
                if constexpr (pre_jacobi(PRE)) dv[u] = dinv[i];
                if constexpr (XP) {
                    xv[u] = x[i];
                    pv[u] = p[i];
                }
                const int el = gx_divf(j, G.n3, G.inv_n3f);
                ge[u] = el;
                sh[u] = gx_issue(gv[u], gc[u], go[u], gs[u], gm[u], w, tabs + el * kGxStride, G,
                                 j - el * G.n3, i);
                if (sh[u]) ps[u] = XP ? pv[u] : p[i];
            }
        }
        if constexpr (XF) {
            const int k = s->iter, lo = s->xdone > k - kRing + 1 ? s->xdone : k - kRing + 1;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = j0 + u * kRedThreads;
                if (j < np) {
                    const int64_t i = beg + j;
                    x[i] = ring_fold<double, FMA>(x[i], ring, rs, i, s, lo, k + 1);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (sh[u]) {
                double v = gx_finish(gv[u], gc[u], go[u], gs[u], vsum + 8 * ge[u],
                                     wv[u]) +
                           0.1 * ps[u];
                if (gm[u]) v = 0.0;
                wv[u] = v;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * kRedThreads;
            if (j >= np) continue;
            const int64_t i = beg + j;
            if constexpr (XP) {
                if constexpr (FMA)
                    x[i] = __builtin_fma(alpha, pv[u], xv[u]);
                else
                    x[i] = xv[u] + alpha * pv[u];
            }
            if constexpr (PRE == kPreJacobiZ) {
                double zi;
                if constexpr (FMA)
                    zi = __builtin_fma(alphm, dv[u] * wv[u], rv[u]);
                else
                    zi = rv[u] + alphm * (dv[u] * wv[u]);
                r[i] = zi;
                const double ri = zi / dv[u];
                const double rc = ri * inv_mult(fv[u]);
                a0 = a0 + rc * ri;
                a1 = a1 + rc * zi;
            } else {
                double ri;
                if constexpr (FMA)
                    ri = __builtin_fma(alphm, wv[u], rv[u]);
                else
                    ri = rv[u] + alphm * wv[u];
                r[i] = ri;
                const double rc = ri * inv_mult(fv[u]);
                a0 = a0 + rc * ri;
                if constexpr (PRE == NEKBONE_PRECOND_JACOBI) a1 = a1 + rc * (ri * dv[u]);
            }
        }
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[blk] = a0;
        part[nblk + blk] = a1;
    }
}

template <int PRE, bool FMA, bool XP, bool XF = false>
__global__ void __launch_bounds__(kRedThreads)
    k_update_gx(double *__restrict__ r, const double *__restrict__ w,
                const double *__restrict__ dinv, const uint8_t *__restrict__ flag,
                double *__restrict__ x, const double *__restrict__ p, int nelt,
                double *__restrict__ part, int nblk, const CgScalars *__restrict__ s, int rev,
                GxView G, const double *__restrict__ ring = nullptr, int64_t rs = 0)
{
    update_gx_body<PRE, FMA, XP, XF>(r, w, dinv, flag, x, p, nelt, part, nblk, s,
                                     sweep_block(blockIdx.x, nblk, rev), G, ring, rs);
}

// x-ring, end of a solve: x receives alpha_j p_j of the iterations j in
// [xdone, iter) that no flush has folded in yet; k_ring_mark then records
// them as folded.
template <typename T, bool FMA>
__global__ void k_ring_tail(T *__restrict__ x, const T *__restrict__ ring, int64_t n,
                            const CgScalars *__restrict__ s)
{
    const int lo = s->xdone, hi = s->iter;
    if (hi <= lo) return;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = ring_fold<T, FMA>(x[i], ring, n, i, s, lo, hi);
}

__global__ void k_ring_mark(CgScalars *__restrict__ s) { s->xdone = s->iter; }

template <int PRE>
__global__ void __launch_bounds__(kFinalThreads)
    k_update_final(const double *__restrict__ part, int nblk, CgScalars *__restrict__ s,
                   double *__restrict__ hist, int hcap, int set_xpending, int xmark = 0)
{
    update_final_body<PRE, kFinalThreads>(part, nblk, s, hist, hcap, set_xpending, xmark);
}

// ---- batched forms: problem q owns workgroups [start[q], start[q+1]) of
// the vector kernels and workgroup q of the scalar kernels ----------------
__device__ __forceinline__ int owner(const int *__restrict__ start, int nprob, int b)
{
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (start[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int PRE, bool FMA, bool XP>
__global__ void __launch_bounds__(kRedThreads)
    k_update_batch(const CgBatchArgs *__restrict__ args, const int *__restrict__ start, int nprob)
{
    const int q = owner(start, nprob, blockIdx.x);
    const CgBatchArgs A = args[q];
    update_body<PRE, FMA, XP>(A.r, A.w, A.dinv, A.flag, A.x, A.p, A.n, A.part_upd, A.nupd,
                              A.scal, blockIdx.x - start[q]);
}

// Batched gather-form update: problem q owns workgroups
// [start[q], start[q+1]), each the workgroup of its single-problem launch.
template <int PRE, bool FMA, bool XP>
__global__ void __launch_bounds__(kRedThreads)
    k_update_gx_batch(const CgBatchArgs *__restrict__ args, const int *__restrict__ start,
                      int nprob)
{
    const int q = owner(start, nprob, blockIdx.x);
    const CgBatchArgs &A = args[q];
    update_gx_body<PRE, FMA, XP, false>(A.r, A.w, A.dinv, A.flag, A.x, A.p, A.nelt, A.part_upd,
                                        A.nupd, A.scal, blockIdx.x - start[q], A.G, nullptr, 0);
}

__global__ void __launch_bounds__(kFinalThreads) k_pap_final_batch(const CgBatchArgs *__restrict__ args)
{
    const CgBatchArgs A = args[blockIdx.x];
    pap_final_body<kFinalThreads>(A.part, A.npap, A.scal, A.hist, A.hcap);
}

template <int PRE>
__global__ void __launch_bounds__(kFinalThreads)
    k_update_final_batch(const CgBatchArgs *__restrict__ args, int set_xpending)
{
    const CgBatchArgs A = args[blockIdx.x];
    update_final_body<PRE, kFinalThreads>(A.part_upd, A.nupd, A.scal, A.hist, A.hcap, set_xpending);
}

// The x update of the last completed iteration (the element kernel of
// the next iteration would have applied it).
template <bool FMA>
__global__ void k_x_finish(double *__restrict__ x, const double *__restrict__ p, int64_t n,
                           const CgScalars *__restrict__ s)
{
    if (!s->xpending) return;
    const double alpha = s->alpha;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (FMA)
        x[i] = __builtin_fma(alpha, p[i], x[i]);
    else
        x[i] = x[i] + alpha * p[i];
}

// ---- generic path (any c) ----------------------------------------------

// p = beta p + z (add2s1, src/math.f:534-541), z = r or r*dinv.
template <int PRE, bool FMA>
__global__ void k_pupdate(double *__restrict__ p, const double *__restrict__ r,
                          const double *__restrict__ dinv, int64_t n,
                          const CgScalars *__restrict__ s)
{
    if (s->done) return;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double z = r[i];
    if constexpr (PRE == NEKBONE_PRECOND_JACOBI) z = z * dinv[i];
    if constexpr (FMA)
        p[i] = __builtin_fma(s->beta, p[i], z);
    else
        p[i] = s->beta * p[i] + z;
}

// x += alpha p; r += (-alpha) w; parts of rtr and rtz with the c array.
template <int PRE, bool FMA>
__global__ void __launch_bounds__(kRedThreads)
    k_update_generic(double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
                     const double *__restrict__ w, const double *__restrict__ c,
                     const double *__restrict__ dinv, int64_t n, double *__restrict__ part,
                     int nblk, const CgScalars *__restrict__ s)
{
    if (s->done) return;
    const double alpha = s->alpha, alphm = -alpha;
    double a0 = 0.0, a1 = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kRedThreads;
    for (int64_t i = int64_t(blockIdx.x) * kRedThreads + threadIdx.x; i < n; i += stride) {
        double xi, ri;
        if constexpr (FMA) {
            xi = __builtin_fma(alpha, p[i], x[i]);
            ri = __builtin_fma(alphm, w[i], r[i]);
        } else {
            xi = x[i] + alpha * p[i];
            ri = r[i] + alphm * w[i];
        }
        x[i] = xi;
        r[i] = ri;
        const double rc = ri * c[i];
        a0 = a0 + rc * ri;
        if constexpr (PRE == NEKBONE_PRECOND_JACOBI) a1 = a1 + rc * (ri * dinv[i]);
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a0;
        part[nblk + blockIdx.x] = a1;
    }
}

template <bool PAP>
__global__ void __launch_bounds__(kRedThreads)
    k_post_cg(double *__restrict__ w, const double *__restrict__ p, const double *__restrict__ c,
              const uint8_t *__restrict__ flag, int mbit, int64_t n, double *__restrict__ part,
              const CgScalars *__restrict__ s)
{
    if (s->done) return;
    double acc = 0.0, dummy = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kRedThreads;
    for (int64_t i = int64_t(blockIdx.x) * kRedThreads + threadIdx.x; i < n; i += stride) {
        double v = w[i] + 0.1 * p[i];
        if (flag[i] & mbit) v = 0.0;
        w[i] = v;
        acc = acc + v * c[i] * p[i];
    }
    block_sum2(acc, dummy);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

int grid1d(int64_t n, int t) { return int((n + t - 1) / t); }

// Partials layout of the fused path:
//   [0, NA)            element-kernel pap parts
//   [NA, NA+NG)        gather-scatter pap parts
//   [NA+NG, +2*NU)     update rtr parts, then rtz parts
int64_t n_ax(const nekbone_problem *P) { return ax_blocks(P); }
int64_t n_gs(const nekbone_problem *P) { return gs_blocks(P); }
int n_upd(const nekbone_problem *P)
{
    int64_t b = (P->n + int64_t(kRedThreads) * 16 - 1) / (int64_t(kRedThreads) * 16);
    return int(std::max<int64_t>(1, std::min<int64_t>(b, 8192)));
}

}  // namespace

int cg_update_blocks(const nekbone_problem *P) { return n_upd(P); }

int launch_pap_final(nekbone_problem *P, const double *part, int64_t npart)
{
    if (int rc = timed_begin(P, 4)) return rc;
    hipLaunchKernelGGL(k_pap_final, dim3(1), dim3(kFinalThreads), 0, stream(), part, npart,
                       P->d_scal, P->d_hist, 0);
    NB_HIP(hipGetLastError());
    return timed_end(P, 4);
}

int launch_update_final(nekbone_problem *P, const double *part, int nu, int precond,
                        int set_xpending, int xmark)
{
    if (int rc = timed_begin(P, 4)) return rc;
    if (precond == NEKBONE_PRECOND_JACOBI)
        hipLaunchKernelGGL(k_update_final<NEKBONE_PRECOND_JACOBI>, dim3(1), dim3(kFinalThreads),
                           0, stream(), part, nu, P->d_scal, P->d_hist, 0, set_xpending, xmark);
    else
        hipLaunchKernelGGL(k_update_final<NEKBONE_PRECOND_IDENTITY>, dim3(1),
                           dim3(kFinalThreads), 0, stream(), part, nu, P->d_scal, P->d_hist, 0,
                           set_xpending, xmark);
    NB_HIP(hipGetLastError());
    return timed_end(P, 4);
}

int64_t cg_partials_needed(const nekbone_problem *P)
{
    // element-kernel parts: at most one per element (smallest workgroup);
    // update parts: the generic update's or the gather form's workgroups
    const int eb = std::min(gx_eb(P->nx1, kGxPts64), gx_eb(P->nx1, kGxPts32));
    const int64_t nux = std::max<int64_t>(n_upd(P), (P->nelt + eb - 1) / eb);
    return std::max<int64_t>(P->nelt + n_gs(P) + 2 * nux, 2 * int64_t(kMaxRedBlocks));
}

bool cg_zform(const nekbone_problem *P, int precond)
{
    return precond == NEKBONE_PRECOND_JACOBI && P->c_is_mult && P->fused;
}

int launch_cg_init(nekbone_problem *P, double *x, double *r, const double *f, int mbit,
                   int precond, double tol2, bool zform, bool fp64_iter)
{
    // every solve decides afresh whether its right-hand side allows the
    // gather form, and with it the x-ring
    if (int rc = cg_gather_check(P, f)) return rc;
    P->ring_on = fp64_iter && cg_gather(P) && ensure_ring(P);
    const int nb = reduce_blocks(P->n);
    if (precond == NEKBONE_PRECOND_JACOBI && zform) {
        hipLaunchKernelGGL(k_cg_init<kPreJacobiZ>, dim3(nb), dim3(kRedThreads), 0, stream(), x,
                           r, f, P->d_c, P->d_dinv, P->d_flag, mbit, P->n, P->d_part, nb);
        hipLaunchKernelGGL(k_cg_init_final<NEKBONE_PRECOND_JACOBI>, dim3(1), dim3(kFinalThreads),
                           0, stream(), P->d_part, nb, P->d_scal, P->d_hist, P->hist_cap, tol2);
    } else if (precond == NEKBONE_PRECOND_JACOBI) {
        hipLaunchKernelGGL(k_cg_init<NEKBONE_PRECOND_JACOBI>, dim3(nb), dim3(kRedThreads), 0,
                           stream(), x, r, f, P->d_c, P->d_dinv, P->d_flag, mbit, P->n,
                           P->d_part, nb);
        hipLaunchKernelGGL(k_cg_init_final<NEKBONE_PRECOND_JACOBI>, dim3(1), dim3(kFinalThreads),
                           0, stream(), P->d_part, nb, P->d_scal, P->d_hist, P->hist_cap, tol2);
    } else {
        hipLaunchKernelGGL(k_cg_init<NEKBONE_PRECOND_IDENTITY>, dim3(nb), dim3(kRedThreads), 0,
                           stream(), x, r, f, P->d_c, P->d_dinv, P->d_flag, mbit, P->n,
                           P->d_part, nb);
        hipLaunchKernelGGL(k_cg_init_final<NEKBONE_PRECOND_IDENTITY>, dim3(1),
                           dim3(kFinalThreads), 0, stream(), P->d_part, nb, P->d_scal, P->d_hist,
                           P->hist_cap, tol2);
    }
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

#define NB_PRE_FMA(FN, ...)                                          \
    do {                                                             \
        if (jac) {                                                   \
            if (fma)                                                 \
                FN(NEKBONE_PRECOND_JACOBI, true, __VA_ARGS__);       \
            else                                                     \
                FN(NEKBONE_PRECOND_JACOBI, false, __VA_ARGS__);      \
        } else {                                                     \
            if (fma)                                                 \
                FN(NEKBONE_PRECOND_IDENTITY, true, __VA_ARGS__);     \
            else                                                     \
                FN(NEKBONE_PRECOND_IDENTITY, false, __VA_ARGS__);    \
        }                                                            \
    } while (0)

// Gather form of the fused iteration (cg_gather): element kernel (pap
// parts over all points) -> pap/alpha -> update kernel that also sums the
// shared points' copies -> rtr/beta.  Two vector passes and two scalar
// kernels; w, and with it x and z, bit-identical to the gather-scatter
// form, pap summed in another order.  The element kernel walks upward and
// the update downward, so each starts where the other ended.
//
// With the x-ring (P->ring_on) the element kernel takes the previous
// direction from ring slot (k-1) % kRing and stores the new one in slot
// k % kRing, without touching x; the update kernel of every kRing-th
// iteration folds the ring's alpha p into x.  zin (HSMG): the element
// kernel reads the preconditioner's z, the update takes r itself.
static int iteration_gather(nekbone_problem *P, double *x, double *r, int precond, int mbit,
                            int k, const double *zin = nullptr)
{
    const bool jac = precond == NEKBONE_PRECOND_JACOBI;
    const bool fma = P->arith == NEKBONE_ARITH_FMA;
    const int64_t na = n_ax(P), ng = n_gs(P);
    const GxView G = gx_view(P, kGxPts64);
    const int nu = int((P->nelt + G.eb - 1) / G.eb);
    double *part_upd = P->d_part + na + ng;
    const bool ring = P->ring_on;
    const int64_t n = P->n;
    double *pcur = ring ? P->d_pring + int64_t(k % kRing) * n : P->d_p;
    double *pprev = ring ? P->d_pring + int64_t((k + kRing - 1) % kRing) * n : P->d_p;
    const bool xf = ring && k % kRing == kRing - 1;
    if (int rc = launch_ax_cg(P, x, zin ? zin : r, mbit, 0, 1, ring ? pprev : nullptr,
                              ring ? pcur : nullptr))
        return rc;
    if (int rc = timed_begin(P, 4)) return rc;
    hipLaunchKernelGGL(k_pap_final, dim3(1), dim3(kFinalThreads), 0, stream(), P->d_part, na,
                       P->d_scal, P->d_hist, P->hist_cap);
    NB_HIP(hipGetLastError());
    if (int rc = timed_end(P, 4)) return rc;
    if (int rc = timed_begin(P, 3)) return rc;
    const bool xp = !ring && !P->axv.xu;
#define NB_UGX3(PRE, F, XP, XF)                                                                   \
    hipLaunchKernelGGL((k_update_gx<PRE, F, XP, XF>), dim3(nu), dim3(kRedThreads), 0, stream(),  \
                       r, P->d_w, P->d_dinv, P->d_flag, x, static_cast<const double *>(pcur),    \
                       P->nelt, part_upd, nu, P->d_scal, 1, G,                                    \
                       static_cast<const double *>(P->d_pring), n)
#define NB_UGX(PRE, F)                         \
    do {                                       \
        if (xf)                                \
            NB_UGX3(PRE, F, false, true);      \
        else if (xp)                           \
            NB_UGX3(PRE, F, true, false);      \
        else                                   \
            NB_UGX3(PRE, F, false, false);     \
    } while (0)
    if (jac) {
        if (fma)
            NB_UGX(kPreJacobiZ, true);
        else
            NB_UGX(kPreJacobiZ, false);
    } else {
        if (fma)
            NB_UGX(NEKBONE_PRECOND_IDENTITY, true);
        else
            NB_UGX(NEKBONE_PRECOND_IDENTITY, false);
    }
#undef NB_UGX
#undef NB_UGX3
    NB_HIP(hipGetLastError());
    if (int rc = timed_end(P, 3)) return rc;
    if (int rc = timed_begin(P, 4)) return rc;
    const int xpend = ring ? 0 : P->axv.xu;
    if (jac)
        hipLaunchKernelGGL(k_update_final<NEKBONE_PRECOND_JACOBI>, dim3(1), dim3(kFinalThreads),
                           0, stream(), part_upd, nu, P->d_scal, P->d_hist, P->hist_cap, xpend,
                           xf ? 1 : 0);
    else
        hipLaunchKernelGGL(k_update_final<NEKBONE_PRECOND_IDENTITY>, dim3(1),
                           dim3(kFinalThreads), 0, stream(), part_upd, nu, P->d_scal, P->d_hist,
                           P->hist_cap, xpend, xf ? 1 : 0);
    NB_HIP(hipGetLastError());
    return timed_end(P, 4);
}

// zin: the vector the element kernel reads in place of r (an explicit
// z = M^-1 r from the HSMG preconditioner); null: r itself.
static int iteration_fused(nekbone_problem *P, double *x, double *r, int precond, int mbit,
                           const double *zin = nullptr, int k = 0)
{
    if (cg_gather(P)) return iteration_gather(P, x, r, precond, mbit, k, zin);
    const bool jac = precond == NEKBONE_PRECOND_JACOBI;
    const bool fma = P->arith == NEKBONE_ARITH_FMA;
    const int64_t na = n_ax(P), ng = n_gs(P);
    const int nu = n_upd(P);
    double *part_upd = P->d_part + na + ng;
    // sweep directions: element kernel one way, gather-scatter back, update
    // forward again; the next iteration starts in the other direction
    const int rev = P->sweep & 1;
    ++P->sweep;
    // Jacobi runs in z form (r holds z = M^-1 r): the element kernel's
    // identity mode reads z alone
    if (int rc = launch_ax_cg(P, x, zin ? zin : r, mbit, rev)) return rc;
    // the element kernel's partials are folded into the gather-scatter
    // kernel's when it has workgroups enough, so the final sum reads fewer
    const bool fold = cg_fold_partials(P);
    if (int rc = launch_gs_cg(P, P->d_w, P->d_p, mbit, P->d_part + na, !rev,
                              fold ? P->d_part : nullptr, fold ? na : 0))
        return rc;
    if (int rc = timed_begin(P, 4)) return rc;
    hipLaunchKernelGGL(k_pap_final, dim3(1), dim3(kFinalThreads), 0, stream(),
                       fold ? P->d_part + na : P->d_part, fold ? ng : na + ng, P->d_scal,
                       P->d_hist, P->hist_cap);
    NB_HIP(hipGetLastError());
    if (int rc = timed_end(P, 4)) return rc;
    if (int rc = timed_begin(P, 3)) return rc;
    const bool xp = !P->axv.xu;  // x update here unless the element kernel carries it
#define NB_UPD2(PRE, F, XP)                                                                  \
    hipLaunchKernelGGL((k_update<PRE, F, XP>), dim3(nu), dim3(kRedThreads), 0, stream(), r,    \
                       P->d_w, P->d_dinv, P->d_flag, x, P->d_p, P->n, part_upd, nu, P->d_scal, \
                       rev)
#define NB_UPD(PRE, F, _)     \
    do {                      \
        if (xp)               \
            NB_UPD2(PRE, F, true); \
        else                  \
            NB_UPD2(PRE, F, false); \
    } while (0)
    if (jac) {
        if (fma)
            NB_UPD(kPreJacobiZ, true, 0);
        else
            NB_UPD(kPreJacobiZ, false, 0);
    } else {
        if (fma)
            NB_UPD(NEKBONE_PRECOND_IDENTITY, true, 0);
        else
            NB_UPD(NEKBONE_PRECOND_IDENTITY, false, 0);
    }
#undef NB_UPD
#undef NB_UPD2
    NB_HIP(hipGetLastError());
    if (int rc = timed_end(P, 3)) return rc;
    if (int rc = timed_begin(P, 4)) return rc;
    if (jac)
        hipLaunchKernelGGL(k_update_final<NEKBONE_PRECOND_JACOBI>, dim3(1), dim3(kFinalThreads),
                           0, stream(), part_upd, nu, P->d_scal, P->d_hist, P->hist_cap,
                           P->axv.xu);
    else
        hipLaunchKernelGGL(k_update_final<NEKBONE_PRECOND_IDENTITY>, dim3(1),
                           dim3(kFinalThreads), 0, stream(), part_upd, nu, P->d_scal, P->d_hist,
                           P->hist_cap, P->axv.xu);
    NB_HIP(hipGetLastError());
    return timed_end(P, 4);
}

static int iteration_generic(nekbone_problem *P, double *x, double *r, int precond, int mbit)
{
    const bool jac = precond == NEKBONE_PRECOND_JACOBI;
    const bool fma = P->arith == NEKBONE_ARITH_FMA;
    const int nb = reduce_blocks(P->n);
#define NB_PUPD(PRE, F, _)                                                                  \
    hipLaunchKernelGGL((k_pupdate<PRE, F>), dim3(grid1d(P->n, 256)), dim3(256), 0, stream(), \
                       P->d_p, r, P->d_dinv, P->n, P->d_scal)
    NB_PRE_FMA(NB_PUPD, 0);
#undef NB_PUPD
    NB_HIP(hipGetLastError());
    if (int rc = launch_ax_local(P, P->d_w, P->d_p)) return rc;
    if (int rc = launch_gs(P, P->d_w, P->d_scal)) return rc;
    hipLaunchKernelGGL(k_post_cg<true>, dim3(nb), dim3(kRedThreads), 0, stream(), P->d_w, P->d_p,
                       P->d_c, P->d_flag, mbit, P->n, P->d_part, P->d_scal);
    hipLaunchKernelGGL(k_pap_final, dim3(1), dim3(kFinalThreads), 0, stream(), P->d_part,
                       int64_t(nb), P->d_scal, P->d_hist, P->hist_cap);
#define NB_UPD(PRE, F, _)                                                                       \
    hipLaunchKernelGGL((k_update_generic<PRE, F>), dim3(nb), dim3(kRedThreads), 0, stream(), x, \
                       r, P->d_p, P->d_w, P->d_c, P->d_dinv, P->n, P->d_part, nb, P->d_scal)
    NB_PRE_FMA(NB_UPD, 0);
#undef NB_UPD
    if (jac)
        hipLaunchKernelGGL(k_update_final<NEKBONE_PRECOND_JACOBI>, dim3(1), dim3(kFinalThreads),
                           0, stream(), P->d_part, nb, P->d_scal, P->d_hist, P->hist_cap, 0);
    else
        hipLaunchKernelGGL(k_update_final<NEKBONE_PRECOND_IDENTITY>, dim3(1),
                           dim3(kFinalThreads), 0, stream(), P->d_part, nb, P->d_scal, P->d_hist,
                           P->hist_cap, 0);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

namespace {

// rtz = glsc3(r, c, z) (src/cg.f:52) with an explicit z
__global__ void __launch_bounds__(kRedThreads)
    k_rtz(const double *__restrict__ r, const double *__restrict__ c,
          const double *__restrict__ z, int64_t n, double *__restrict__ part,
          const CgScalars *__restrict__ s)
{
    if (s->done) return;
    double a0 = 0.0, dummy = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kRedThreads;
    for (int64_t i = int64_t(blockIdx.x) * kRedThreads + threadIdx.x; i < n; i += stride)
        a0 = a0 + r[i] * c[i] * z[i];
    block_sum2(a0, dummy);
    if (threadIdx.x == 0) part[blockIdx.x] = a0;
}

// rtz1 and beta of this iteration (src/cg.f:52-55) from the rtz parts;
// replaces what the previous update's bookkeeping assumed (z = r).
__global__ void __launch_bounds__(kFinalThreads)
    k_rtz_final(const double *__restrict__ part, int nblk, CgScalars *__restrict__ s,
                double *__restrict__ hist, int hcap)
{
    if (s->done) return;
    double rtz, dummy;
    sum_region2<kFinalThreads>(part, nblk, nullptr, 0, rtz, dummy);
    if (threadIdx.x == 0) {
        s->rtz1 = rtz;
        s->beta = (s->iter == 0) ? 0.0 : rtz / s->rtz2;
        const int h = 1 + 3 * s->iter;
        if (h < hcap) hist[h] = rtz;
    }
}

// p = beta p + z (add2s1, src/cg.f:56)
template <bool FMA>
__global__ void k_pupdate_z(double *__restrict__ p, const double *__restrict__ z, int64_t n,
                            const CgScalars *__restrict__ s)
{
    if (s->done) return;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (FMA)
        p[i] = __builtin_fma(s->beta, p[i], z[i]);
    else
        p[i] = s->beta * p[i] + z[i];
}

}  // namespace

int launch_cg_iteration_z(nekbone_problem *P, double *x, double *r, const double *z, int mbit,
                          int k, int rtz_parts)
{
    const bool fma = P->arith == NEKBONE_ARITH_FMA;
    const int nb = reduce_blocks(P->n);
    if (rtz_parts <= 0)
        hipLaunchKernelGGL(k_rtz, dim3(nb), dim3(kRedThreads), 0, stream(),
                           static_cast<const double *>(r), static_cast<const double *>(P->d_c),
                           z, P->n, P->d_part, static_cast<const CgScalars *>(P->d_scal));
    hipLaunchKernelGGL(k_rtz_final, dim3(1), dim3(kFinalThreads), 0, stream(),
                       static_cast<const double *>(P->d_part), rtz_parts > 0 ? rtz_parts : nb,
                       P->d_scal, P->d_hist, P->hist_cap);
    NB_HIP(hipGetLastError());
    // with c = 1/multiplicity the rest is the fused identity iteration with
    // the element kernel reading z (p = z + beta p fused into it)
    if (P->c_is_mult && P->fused)
        return iteration_fused(P, x, r, NEKBONE_PRECOND_IDENTITY, mbit, z, k);
    if (fma)
        hipLaunchKernelGGL(k_pupdate_z<true>, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(),
                           P->d_p, z, P->n, static_cast<const CgScalars *>(P->d_scal));
    else
        hipLaunchKernelGGL(k_pupdate_z<false>, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(),
                           P->d_p, z, P->n, static_cast<const CgScalars *>(P->d_scal));
    NB_HIP(hipGetLastError());
    if (int rc = launch_ax_local(P, P->d_w, P->d_p)) return rc;
    if (int rc = launch_gs(P, P->d_w, P->d_scal)) return rc;
    hipLaunchKernelGGL(k_post_cg<true>, dim3(nb), dim3(kRedThreads), 0, stream(), P->d_w, P->d_p,
                       P->d_c, P->d_flag, mbit, P->n, P->d_part, P->d_scal);
    hipLaunchKernelGGL(k_pap_final, dim3(1), dim3(kFinalThreads), 0, stream(), P->d_part,
                       int64_t(nb), P->d_scal, P->d_hist, P->hist_cap);
    if (fma)
        hipLaunchKernelGGL((k_update_generic<NEKBONE_PRECOND_IDENTITY, true>), dim3(nb),
                           dim3(kRedThreads), 0, stream(), x, r, P->d_p, P->d_w, P->d_c,
                           P->d_dinv, P->n, P->d_part, nb, P->d_scal);
    else
        hipLaunchKernelGGL((k_update_generic<NEKBONE_PRECOND_IDENTITY, false>), dim3(nb),
                           dim3(kRedThreads), 0, stream(), x, r, P->d_p, P->d_w, P->d_c,
                           P->d_dinv, P->n, P->d_part, nb, P->d_scal);
    hipLaunchKernelGGL(k_update_final<NEKBONE_PRECOND_IDENTITY>, dim3(1), dim3(kFinalThreads), 0,
                       stream(), P->d_part, nb, P->d_scal, P->d_hist, P->hist_cap, 0);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

// One PCG iteration (src/cg.f:48-75), no host round trip; k is its index
// within the solve.
int launch_cg_iteration(nekbone_problem *P, double *x, double *r, int precond, int mbit, int k)
{
    if (P->c_is_mult && P->fused) return iteration_fused(P, x, r, precond, mbit, nullptr, k);
    return iteration_generic(P, x, r, precond, mbit);
}

CgBatchArgs cg_batch_args(nekbone_problem *P, double *x, double *r, bool gx)
{
    CgBatchArgs A{};
    if (gx) {
        // gather form: the element kernel's parts alone make pap; the
        // update parts come from the gather-form update's workgroups
        A.r = r;
        A.w = P->d_w;
        A.dinv = P->d_dinv;
        A.flag = P->d_flag;
        A.x = x;
        A.p = P->d_p;
        A.n = P->n;
        A.part = P->d_part;
        A.npap = n_ax(P);
        A.part_upd = P->d_part + n_ax(P) + n_gs(P);
        A.G = gx_view(P, kGxPts64);
        A.nupd = int((P->nelt + A.G.eb - 1) / A.G.eb);
        A.nelt = P->nelt;
        A.scal = P->d_scal;
        A.hist = P->d_hist;
        A.hcap = P->hist_cap;
        return A;
    }
    A.r = r;
    A.w = P->d_w;
    A.dinv = P->d_dinv;
    A.flag = P->d_flag;
    A.x = x;
    A.p = P->d_p;
    A.n = P->n;
    const bool fold = cg_fold_partials(P);  // as in the single-problem iteration
    A.part = fold ? P->d_part + n_ax(P) : P->d_part;
    A.npap = fold ? n_gs(P) : n_ax(P) + n_gs(P);
    A.part_upd = P->d_part + n_ax(P) + n_gs(P);
    A.nupd = n_upd(P);
    A.scal = P->d_scal;
    A.hist = P->d_hist;
    A.hcap = P->hist_cap;
    return A;
}

// The scalar and vector-update kernels of one fused iteration for a batch
// of problems (the element and gather-scatter kernels are launched per
// order group before): pap/alpha, r update with rtr and rtz, rtr/beta.
int launch_cg_batch_tail(const CgBatchArgs *d_args, const int *d_ustart, int nprob, int nupd_total,
                         int precond, bool fma, bool xp, int set_xpending, bool gx)
{
    hipLaunchKernelGGL(k_pap_final_batch, dim3(nprob), dim3(kFinalThreads), 0, stream(), d_args);
    if (gx) {
#define NB_UGB(PRE, F, X)                                                                     \
    hipLaunchKernelGGL((k_update_gx_batch<PRE, F, X>), dim3(nupd_total), dim3(kRedThreads), 0, \
                       stream(), d_args, d_ustart, nprob)
        const bool j = precond == NEKBONE_PRECOND_JACOBI;
        if (j) {
            if (fma) {
                if (xp) NB_UGB(kPreJacobiZ, true, true); else NB_UGB(kPreJacobiZ, true, false);
            } else {
                if (xp) NB_UGB(kPreJacobiZ, false, true); else NB_UGB(kPreJacobiZ, false, false);
            }
        } else {
            if (fma) {
                if (xp) NB_UGB(NEKBONE_PRECOND_IDENTITY, true, true); else NB_UGB(NEKBONE_PRECOND_IDENTITY, true, false);
            } else {
                if (xp) NB_UGB(NEKBONE_PRECOND_IDENTITY, false, true); else NB_UGB(NEKBONE_PRECOND_IDENTITY, false, false);
            }
        }
#undef NB_UGB
        NB_HIP(hipGetLastError());
        if (j)
            hipLaunchKernelGGL(k_update_final_batch<NEKBONE_PRECOND_JACOBI>, dim3(nprob),
                               dim3(kFinalThreads), 0, stream(), d_args, set_xpending);
        else
            hipLaunchKernelGGL(k_update_final_batch<NEKBONE_PRECOND_IDENTITY>, dim3(nprob),
                               dim3(kFinalThreads), 0, stream(), d_args, set_xpending);
        NB_HIP(hipGetLastError());
        return NEKBONE_OK;
    }
#define NB_UB(PRE, F, X)                                                                     \
    hipLaunchKernelGGL((k_update_batch<PRE, F, X>), dim3(nupd_total), dim3(kRedThreads), 0, \
                       stream(), d_args, d_ustart, nprob)
    const bool jac = precond == NEKBONE_PRECOND_JACOBI;
    if (jac) {
        if (fma) {
            if (xp) NB_UB(kPreJacobiZ, true, true); else NB_UB(kPreJacobiZ, true, false);
        } else {
            if (xp) NB_UB(kPreJacobiZ, false, true); else NB_UB(kPreJacobiZ, false, false);
        }
    } else {
        if (fma) {
            if (xp) NB_UB(NEKBONE_PRECOND_IDENTITY, true, true); else NB_UB(NEKBONE_PRECOND_IDENTITY, true, false);
        } else {
            if (xp) NB_UB(NEKBONE_PRECOND_IDENTITY, false, true); else NB_UB(NEKBONE_PRECOND_IDENTITY, false, false);
        }
    }
#undef NB_UB
    if (jac)
        hipLaunchKernelGGL(k_update_final_batch<NEKBONE_PRECOND_JACOBI>, dim3(nprob),
                           dim3(kFinalThreads), 0, stream(), d_args, set_xpending);
    else
        hipLaunchKernelGGL(k_update_final_batch<NEKBONE_PRECOND_IDENTITY>, dim3(nprob),
                           dim3(kFinalThreads), 0, stream(), d_args, set_xpending);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int cg_ax_blocks(const nekbone_problem *P) { return int(n_ax(P)); }
bool cg_fold_partials(const nekbone_problem *P) { return n_gs(P) * kRedThreads >= n_ax(P); }
int cg_gs_blocks(const nekbone_problem *P) { return int(n_gs(P)); }

static bool gather_applies(const nekbone_problem *P)
{
    return P->gx_ok && P->gather && P->c_is_mult && P->fused &&
           P->mask_mode == NEKBONE_MASK_DIRICHLET;
}

bool cg_gather(const nekbone_problem *P) { return gather_applies(P) && P->gx_cont; }

namespace {
// Continuity of f at the unmasked shared points: every member of a
// point's group holds the same value (else *bad = 1).
__global__ void __launch_bounds__(kRedThreads)
    k_gx_check(const double *__restrict__ f, int nelt, GxView G, int *__restrict__ bad)
{
    __shared__ int32_t tabs[kGxMaxEb * kGxStride];
    const int64_t e0 = int64_t(blockIdx.x) * G.eb;
    const int ne = int(nelt - e0 < G.eb ? nelt - e0 : G.eb);
    gx_rows<kRedThreads>(tabs, G, e0, ne);
    const int np = ne * G.n3;
    const int64_t beg = e0 * G.n3;
    bool ok = true;
    for (int j = threadIdx.x; j < np; j += kRedThreads) {
        const int el = gx_divf(j, G.n3, G.inv_n3f), d = j - el * G.n3;
        int sb, K, o, bit;
        if (!gx_entity(d, G, sb, K, o, bit)) continue;
        const int32_t *tb = tabs + el * kGxStride;
        if (tb[sb] < 0 || ((tb[kGxMask] >> bit) & 1)) continue;
        const double v = f[beg + j];
        for (int m = 0; m < K; ++m) {
            const int s = tb[sb + m];
            if (s < 0) break;
            ok = ok && f[s + o] == v;
        }
    }
    if (!ok) *bad = 1;
}
}  // namespace

int cg_gather_check(nekbone_problem *P, const double *f)
{
    P->gx_cont = true;
    if (!gather_applies(P)) return NEKBONE_OK;
    if (int rc = dzero(P->d_gxbad, sizeof(int))) return rc;
    const GxView G = gx_view(P, kGxPts64);
    const int nb = int((P->nelt + G.eb - 1) / G.eb);
    hipLaunchKernelGGL(k_gx_check, dim3(nb), dim3(kRedThreads), 0, stream(), f, P->nelt, G,
                       P->d_gxbad);
    NB_HIP(hipGetLastError());
    int bad = 0;
    if (int rc = d2h(&bad, P->d_gxbad, sizeof(int))) return rc;
    P->gx_cont = bad == 0;
    return NEKBONE_OK;
}

GxView gx_view(const nekbone_problem *P, int pts)
{
    GxView G{};
    G.tab = P->d_gx;
    G.nx = P->nx1;
    G.I = P->nx1 - 2;
    G.n3 = P->nx1 * P->nx1 * P->nx1;
    G.inv_n3 = 1.0 / double(G.n3);
    G.inv_n3f = 1.0f / float(G.n3);
    G.inv_I = G.I > 0 ? 1.0f / float(G.I) : 0.0f;
    G.inv_I2 = G.I > 0 ? 1.0f / float(G.I * G.I) : 0.0f;
    G.eb = gx_eb(P->nx1, pts);
    return G;
}

// z form: r = z / dinv in place.
__global__ void k_z_to_r(double *__restrict__ r, const double *__restrict__ dinv, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) r[i] = r[i] / dinv[i];
}

int launch_ring_tail(nekbone_problem *P, double *x)
{
    if (P->arith == NEKBONE_ARITH_FMA)
        hipLaunchKernelGGL((k_ring_tail<double, true>), dim3(grid1d(P->n, 256)), dim3(256), 0,
                           stream(), x, static_cast<const double *>(P->d_pring), P->n,
                           static_cast<const CgScalars *>(P->d_scal));
    else
        hipLaunchKernelGGL((k_ring_tail<double, false>), dim3(grid1d(P->n, 256)), dim3(256), 0,
                           stream(), x, static_cast<const double *>(P->d_pring), P->n,
                           static_cast<const CgScalars *>(P->d_scal));
    hipLaunchKernelGGL(k_ring_mark, dim3(1), dim3(1), 0, stream(), P->d_scal);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

bool ensure_ring(nekbone_problem *P)
{
    if (!P->xring) return false;
    if (P->d_pring) return true;
    // keep the ring only where it leaves room: it takes kRing vectors
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) return false;
    const size_t need = sizeof(double) * size_t(kRing) * size_t(P->n);
    if (need + (size_t(4) << 30) > freeb) return false;
    if (dalloc(&P->d_pring, int64_t(kRing) * P->n) != NEKBONE_OK) {
        P->d_pring = nullptr;
        return false;
    }
    return true;
}

int launch_cg_finish(nekbone_problem *P, double *x, double *r, bool zform)
{
    if (zform) {
        hipLaunchKernelGGL(k_z_to_r, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(), r,
                           static_cast<const double *>(P->d_dinv), P->n);
        NB_HIP(hipGetLastError());
    }
    if (P->ring_on) return launch_ring_tail(P, x);
    if (!(P->c_is_mult && P->fused) || !P->axv.xu) return NEKBONE_OK;
    if (P->arith == NEKBONE_ARITH_FMA)
        hipLaunchKernelGGL(k_x_finish<true>, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(), x,
                           P->d_p, P->n, P->d_scal);
    else
        hipLaunchKernelGGL(k_x_finish<false>, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(), x,
                           P->d_p, P->n, P->d_scal);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

}  // namespace nb
