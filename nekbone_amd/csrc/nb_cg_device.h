// Device helpers of the CG kernels: fixed-order reductions of partial
// sums and the scalar bookkeeping of an iteration (alpha, beta, rtr,
// convergence), shared by the single-problem and the batched kernels and
// by the mixed-precision solver.
#pragma once

#include "nb_device.h"

namespace nb {

constexpr int kFinalThreads = 1024;

// Points per thread whose loads the gather-form update kernels keep in
// flight together (measured best of 1, 2, 4, 8 for FP64 and FP32).
constexpr int kGxU = 2;

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

// x / D for 0 <= x < 2^24 and D >= 1, inv = 1/D rounded to float: the
// float quotient is within one of the true one, which the remainder
// corrects.
__device__ __forceinline__ int gx_divf(int x, int D, float inv)
{
    int q = int(float(x) * inv);
    const int rem = x - q * D;
    if (rem < 0)
        --q;
    else if (rem >= D)
        ++q;
    return q;
}

// Entity of in-element device position d (GxView layout): slot base sb,
// slots K, offset o within the entity, mask bit; false for the interior.
__device__ __forceinline__ bool gx_entity(int d, const GxView &G, int &sb, int &K, int &o,
                                          int &bit)
{
    const int I = G.I;
    if (d < 8) {
        sb = 8 * d, K = 8, o = 0, bit = d;
        return true;
    }
    if (d < 8 + 12 * I) {
        const int x = d - 8, k = gx_divf(x, I, G.inv_I);
        o = x - k * I, sb = 64 + 4 * k, K = 4, bit = 8 + k;
        return true;
    }
    const int x = d - 8 - 12 * I, I2 = I * I;
    if (x >= 6 * I2) return false;
    const int f = gx_divf(x, I2, G.inv_I2);
    o = x - f * I2, sb = 112 + 2 * f, K = 2, bit = 20 + f;
    return true;
}

// Per-element gather-form update kernels: the member sum of a shared
// point, in two phases so that the loads of all of a thread's points are
// in flight together with the points' own loads.  gx_issue decides from
// the entity table alone whether in-element position d is shared (a
// surface entity with a slot list), loads the members of a face or edge
// point (at most four; the point's own value is not loaded again) into v,
// and for a vertex point (up to eight members) leaves the loads to
// gx_finish, which sums in member order.  tb: the LDS row of the point's
// element.  The state is plain scalars and arrays per point so that it
// stays in registers.
template <typename T>
__device__ __forceinline__ bool gx_issue(T (&v)[4], int &cnt, int &own, int &sb, bool &masked,
                                         const T *__restrict__ w, const int32_t *tb,
                                         const GxView &G, int d, int64_t i)
{
    int K, o, bit;
    if (!gx_entity(d, G, sb, K, o, bit)) return false;
    if (tb[sb] < 0) return false;  // no group: not shared
    masked = (tb[kGxMask] >> bit) & 1;
    cnt = 0;
    own = 0;
    if (K == 8) return true;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int s = m < K ? tb[sb + m] : -1;
        if (s >= 0) {
            ++cnt;
            if (int64_t(s) + o == i)
                own |= 1 << m;
            else
                v[m] = w[s + o];
        }
    }
    return true;
}

// vsum: the member sums of the workgroup's vertex points (gx_vertices).
template <typename T>
__device__ __forceinline__ T gx_finish(const T (&v)[4], int cnt, int own, int sb,
                                       const T *vsum, T self)
{
    if (cnt > 0) {
        T t = (own & 1) ? self : v[0];
#pragma unroll
        for (int m = 1; m < 4; ++m)
            if (m < cnt) t += ((own >> m) & 1) ? self : v[m];
        return t;
    }
    return vsum[sb >> 3];
}

// Member sums of the shared vertex points of the workgroup's ne elements
// (vsum[8 el + v], in member order), one thread per vertex with all its
// member loads in flight; tabs: the elements' table rows in LDS.  Ends
// with a barrier.
template <typename T, int NT>
__device__ __forceinline__ void gx_vertices(T *vsum, const T *__restrict__ w, const int32_t *tabs,
                                            int ne)
{
    for (int t = threadIdx.x; t < 8 * ne; t += NT) {
        const int32_t *sl = tabs + (t >> 3) * kGxStride + 8 * (t & 7);
        T v[8];
#pragma unroll
        for (int m = 0; m < 8; ++m)
            if (sl[m] >= 0) v[m] = w[sl[m]];
        T s = v[0];
#pragma unroll
        for (int m = 1; m < 8; ++m)
            if (sl[m] >= 0) s += v[m];
        vsum[t] = s;
    }
    __syncthreads();
}

// x-ring: x += alpha_j p_j for the iterations j in [lo, hi) (at most
// kRing, the directions in ring slots j % kRing, rs apart), in iteration
// order and with the operations of the per-iteration update, so x comes
// out as it would have step by step.
__device__ __forceinline__ float ring_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double ring_fma(double a, double b, double c)
{
    return __builtin_fma(a, b, c);
}

template <typename T, bool FMA>
__device__ __forceinline__ T ring_fold(T xv, const T *__restrict__ ring, int64_t rs, int64_t i,
                                       const CgScalars *__restrict__ s, int lo, int hi)
{
    T pv[kRing];
#pragma unroll
    for (int t = 0; t < kRing; ++t)
        if (lo + t < hi) pv[t] = ring[int64_t((lo + t) % kRing) * rs + i];
#pragma unroll
    for (int t = 0; t < kRing; ++t) {
        if (lo + t < hi) {
            const T a = T(s->aring[(lo + t) % kRing]);
            if constexpr (FMA)
                xv = ring_fma(a, pv[t], xv);
            else
                xv = xv + a * pv[t];
        }
    }
    return xv;
}

// Mixed-precision CG with device-driven refreshes: whether the FP64
// residual refresh queued after a chunk of inner iterations runs (inner
// recurrence converged, residual reduced by delta since the last refresh,
// or the iteration limit reached), until the FP64 residual has met the
// tolerance.  The refresh kernels evaluate it on scalars that only the
// last of them changes.
__device__ __forceinline__ bool mx_want(const CgScalars *s)
{
    return !s->status && !s->fin &&
           (s->done || s->rtr <= s->delta2 * s->rtr_upd || s->iter >= s->maxit);
}

// Table rows of elements e0 .. e0+ne-1 into LDS (the workgroup's threads).
template <int NT>
__device__ __forceinline__ void gx_rows(int32_t *tabs, const GxView &G, int64_t e0, int ne)
{
    for (int t = threadIdx.x; t < ne * kGxStride; t += NT) tabs[t] = G.tab[e0 * kGxStride + t];
    __syncthreads();
}

// Fixed-order sum of a strided slice with sixteen interleaved accumulators
// (sixteen loads in flight per thread: the single-block final kernels sum
// up to ~60k partials, and their time is the memory latency of the rounds
// of loads), combined pairwise in a fixed tree.
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

// The same for two slices of equal length, the loads of both in flight
// together (sixteen per slice).
template <int NT>
__device__ __forceinline__ void slice_sum_pair(const double *__restrict__ a,
                                               const double *__restrict__ b, int64_t n,
                                               double &sa, double &sb)
{
    constexpr int U = 16;
    double s[U], t[U];
#pragma unroll
    for (int u = 0; u < U; ++u) s[u] = t[u] = 0.0;
    int64_t i = threadIdx.x;
    for (; i + (U - 1) * NT < n; i += U * NT) {
        double va[U], vb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            va[u] = a[i + u * NT];
            vb[u] = b[i + u * NT];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            s[u] = s[u] + va[u];
            t[u] = t[u] + vb[u];
        }
    }
    for (int u = 0; i < n; i += NT, ++u) {
        s[u] = s[u] + a[i];
        t[u] = t[u] + b[i];
    }
#pragma unroll
    for (int h = U / 2; h > 0; h >>= 1)
#pragma unroll
        for (int u = 0; u < h; ++u) {
            s[u] = s[u] + s[u + h];
            t[u] = t[u] + t[u + h];
        }
    sa = s[0];
    sb = t[0];
}

// Sum of a partials region in a fixed order by one block (two regions of
// equal length: their loads interleaved).  (Thirty-two loads in flight for
// one region ran the single-block kernel at 22 instead of 9 us, and so did
// each thread summing a contiguous chunk of 16 or 32, whose loads no
// longer coalesce.)
template <int NT>
__device__ __forceinline__ void sum_region2(const double *__restrict__ a, int64_t na,
                                            const double *__restrict__ b, int64_t nb, double &s0,
                                            double &s1)
{
    double a0, a1;
    if (b && na == nb) {
        slice_sum_pair<NT>(a, b, na, a0, a1);
    } else {
        a0 = slice_sum<NT>(a, na);
        a1 = b ? slice_sum<NT>(b, nb) : 0.0;
    }
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
        s->aring[s->iter % kRing] = s->alpha;
        s->xpending = 0;  // the element kernel of this iteration applied it
        if (!(pap > 0.0) || !isfinite(pap)) {
            s->status = 1;
            if (s->tol2 > 0.0) s->done = 1;
        }
        const int h = 2 + 3 * s->iter;
        if (h < hcap) hist[h] = pap;
    }
}

// xmark: this iteration's update kernel folded the x-ring into x
template <int PRE, int NT>
__device__ __forceinline__ void update_final_body(const double *__restrict__ part, int nblk,
                                                  CgScalars *__restrict__ s,
                                                  double *__restrict__ hist, int hcap,
                                                  int set_xpending, int xmark = 0)
{
    if (s->done) return;
    double rtr, rtz;
    sum_region2<NT>(part, nblk, part + nblk, nblk, rtr, rtz);
    if (threadIdx.x == 0) {
        s->rtr = rtr;
        const int h = 3 + 3 * s->iter;
        if (h < hcap) hist[h] = rtr;
        s->iter += 1;
        if (xmark) s->xdone = s->iter;
        if (set_xpending) s->xpending = 1;
        s->rtz_next = pre_jacobi(PRE) ? rtz : rtr;
        if (s->tol2 > 0.0 && rtr <= s->tol2 * s->rtr0) s->done = 1;
        prepare_next(s, hist, hcap);
    }
}

}  // namespace nb
