// Mixed-precision CG: FP32 inner iterations with FP64 reliable updates.
// There is no counterpart in the reference, whose CG (src/cg.f:2-84) is
// FP64 throughout; the solve stops on the same criterion as
// nekbone_cg_tol_d, evaluated on the FP64 residual of the FP64 operator.
//
// Algorithm (reliable updates, Clark et al., "Solving lattice QCD systems
// of equations using mixed precision solvers on GPUs", 2010, applied to
// Jacobi-preconditioned CG):
//   x64 = 0, r64 = mask(f), rtr0 = (r64, c r64)          FP64
//   r32 = r64, x32 = 0, p32 = 0
//   repeat  (FP32 vectors and operator; dot products accumulated in FP64)
//     z = M^-1 r32, beta = rtz/rtz_prev, p32 = z + beta p32
//     w32 = A p32, alpha = rtz / (w32, c p32)
//     x32 += alpha p32, r32 -= alpha w32
//     when (r32, c r32) <= delta^2 * rtr_upd, or it claims convergence:
//       x64 += x32, x32 = 0
//       r64 = mask(f - A x64) with the FP64 operator, r32 = r64,
//       rtr_upd = (r64, c r64); rtz recomputed from r64 (p32 is kept)
//       stop once rtr_upd <= tol^2 rtr0
// The inner iteration runs the fused kernel sequence of the FP64 solver
// instantiated for float: element kernel (p update, A_L, x update one
// iteration late), gather-scatter with shift/mask/pap, and the update of
// r with both dot products, all without a host round trip.  The host
// checks the device scalars once per chunk of iterations.
#include <algorithm>
#include <cstring>
#include <tuple>

#include "nb_ax_kernel.h"
#include "nb_cg_device.h"

namespace nb {
namespace {

using axk::AxArgsT;
using axk::AxFn32;
using axk::Key32;
using axk::Table32;

Table32 &table32()
{
    static Table32 t;
    static bool done = false;
    if (!done) {
        done = true;
        axk::reg32_low(t);
        axk::reg32_high(t);
        axk::reg32_mf_all(t);
    }
    return t;
}

AxFn32 lookup32(int nx, int epb, int mode, int pf, int xu, int g16 = 0)
{
    auto it = table32().find(Key32(nx, epb, mode, pf, mode == 0 ? 0 : xu, g16));
    return it == table32().end() ? nullptr : it->second;
}

constexpr int kThreads = 256;
int grid(int64_t n) { return int((n + kThreads - 1) / kThreads); }

__global__ void k_to_f32(float *__restrict__ dst, const double *__restrict__ src, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = float(src[i]);
}

// Jacobi diagonal for the inner iterations: FP64 -> FP32 -> bfloat16 bits,
// round to nearest even (the values are positive and finite).
__global__ void k_to_bf16(uint16_t *__restrict__ dst, const double *__restrict__ src,
                          const uint8_t *__restrict__ flag, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = __float_as_uint(float(src[i]));
    if (!flag) {
        dst[i] = uint16_t((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
        return;
    }
    // multiplicity form: five mantissa bits (round to nearest even), the
    // last two bits of the bfloat16 word log2 of the point's multiplicity
    // (1, 2, 4 or 8), so the vector update needs no flag byte; every use
    // of the diagonal reads the same perturbed value (relative change
    // < 2^-5), so the preconditioner stays one fixed diagonal
    const uint32_t t = (b + 0x1FFFFu + ((b >> 18) & 1u)) >> 18;
    const int m = flag[i] >> kMultShift;
    const uint32_t code = m >= 8 ? 3u : m >= 4 ? 2u : m >= 2 ? 1u : 0u;
    dst[i] = uint16_t((t << 2) | code);
}

// FP16 geometric factors: one workgroup per (factor, element) block of
// the FP64 factor planes (block c*nelt + e) finds the largest magnitude,
// stores it as the scale, and writes the factor divided by it, rounded to
// half precision (relative error <= 2^-11 for every value above 2^-14 of
// the largest).
__global__ void __launch_bounds__(256)
    k_to_g16(__half *__restrict__ gh, float *__restrict__ gsc, const double *__restrict__ g,
             int n3, int64_t nelt)
{
    const int64_t base = int64_t(blockIdx.x) * n3;
    double m = 0.0;
    for (int q = threadIdx.x; q < n3; q += blockDim.x) m = fmax(m, fabs(g[base + q]));
    __shared__ double sm[256];
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + h]);
        __syncthreads();
    }
    const double scale = sm[0] > 0.0 ? sm[0] : 1.0;
    if (threadIdx.x == 0) gsc[blockIdx.x] = float(scale);
    const double inv = 1.0 / double(float(scale));
    // six factors per point: factor c of point q of element e at
    // gh[6*(e*n3 + q) + c]
    const int64_t c = blockIdx.x / nelt, e = blockIdx.x - c * nelt;
    for (int q = threadIdx.x; q < n3; q += blockDim.x)
        gh[6 * (e * n3 + q) + c] = __float2half(float(g[base + q] * inv));
}

// FP32 factors for the inner iterations, six per point:
// dst[6*l + c] = src[c*n + l] (l = e*n3 + q).
__global__ void k_g_to_f32_points(float *__restrict__ dst, const double *__restrict__ src,
                                  int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 6 * n) return;
    const int64_t l = i / 6, c = i - l * 6;
    dst[i] = float(src[c * n + l]);
}

// Reference point order <-> device order for float vectors (perm gives
// the device position of each reference point within an element).
template <bool TO_DEV>
__global__ void k_perm32(float *__restrict__ dst, const float *__restrict__ src,
                         const int32_t *__restrict__ perm, int n3, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t e = i / n3;
    const int64_t d = e * n3 + perm[i - e * n3];
    if constexpr (TO_DEV)
        dst[d] = src[i];
    else
        dst[i] = src[d];
}

// Start of the inner iterations: r32 = r64 (with the Jacobi diagonal d:
// z32 = d r64, the vector the inner iterations carry), x32 = p32 = 0; the
// FP64 residual of the initial guess is the first reliable residual.
__global__ void k_mx_start(float *__restrict__ r32, float *__restrict__ x32,
                           float *__restrict__ p32, const double *__restrict__ r64, int64_t n,
                           CgScalars *__restrict__ s, const uint16_t *__restrict__ dinv,
                           double delta2, int maxit)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0) {
        s->rtr_upd = s->rtr;
        s->delta2 = delta2;
        s->maxit = maxit;
        s->fin = 0;
        s->nupd = 0;
    }
    if (i >= n) return;
    r32[i] = dinv ? float(r64[i] * double(axk::dinv_at(dinv, i))) : float(r64[i]);
    x32[i] = 0.0f;
    p32[i] = 0.0f;
}

// r32 -= alpha w32; [XP] x32 += alpha p32; parts of rtr = (r c r) and of
// rtz = (r c z), z = M^-1 r, accumulated in FP64.
// [ZF] (Jacobi) the vector is z = M^-1 r instead of r: z -= alpha (d w),
// and the dot products take r = z / d (d the diagonal as stored), so the
// element kernel reads z alone instead of r and the diagonal.
template <int PRE, bool XP, bool MB = false, bool ZF = false>
__global__ void __launch_bounds__(kRedThreads)
    k_mx_update(float *__restrict__ r, const float *__restrict__ w,
                const uint16_t *__restrict__ dinv, const uint8_t *__restrict__ flag,
                float *__restrict__ x, const float *__restrict__ p, int64_t n,
                double *__restrict__ part, int nblk, const CgScalars *__restrict__ s)
{
    if (s->done) return;
    const float alpha = float(s->alpha);
    double a0 = 0.0, a1 = 0.0;
    constexpr int U = 4;
    const int64_t stride = int64_t(gridDim.x) * kRedThreads * U;
    for (int64_t i0 = int64_t(blockIdx.x) * kRedThreads * U + threadIdx.x; i0 < n; i0 += stride) {
        float rv[U], wv[U], dv[U], xv[U], pv[U];
        uint8_t fv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + int64_t(u) * kRedThreads;
            if (i < n) {
                rv[u] = r[i];
                wv[u] = w[i];
                if constexpr (MB) {
                    // multiplicity form of the diagonal (k_to_bf16): c from its last bits
                    const uint32_t h = dinv[i];
                    dv[u] = __uint_as_float(h << 16);
                    fv[u] = uint8_t(1u << (h & 3u)) << kMultShift;
                } else {
                    fv[u] = flag[i];
                    if constexpr (PRE == NEKBONE_PRECOND_JACOBI) dv[u] = axk::dinv_at(dinv, i);
                }
                if constexpr (XP) {
                    xv[u] = x[i];
                    pv[u] = p[i];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + int64_t(u) * kRedThreads;
            if (i < n) {
                if constexpr (XP) x[i] = __builtin_fmaf(alpha, pv[u], xv[u]);
                if constexpr (ZF) {
                    const float zi = __builtin_fmaf(-alpha, dv[u] * wv[u], rv[u]);
                    r[i] = zi;
                    const double rd = double(zi / dv[u]);
                    const double rc = rd * inv_mult(fv[u]);
                    a0 = a0 + rc * rd;
                    a1 = a1 + rc * double(zi);
                } else {
                    const float ri = __builtin_fmaf(-alpha, wv[u], rv[u]);
                    r[i] = ri;
                    const double rd = double(ri);
                    const double rc = rd * inv_mult(fv[u]);
                    a0 = a0 + rc * rd;
                    if constexpr (PRE == NEKBONE_PRECOND_JACOBI)
                        a1 = a1 + rc * double(ri * dv[u]);
                }
            }
        }
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a0;
        part[nblk + blockIdx.x] = a1;
    }
}

// Gather form of k_mx_update (see cg_gather), one workgroup per G.eb
// consecutive elements with their entity-table rows in LDS: the shared
// points' w is formed here as the sum of their copies in member order,
// + 0.1 p, masked (the mask bit of the entity), exactly as the
// gather-scatter kernel stores it; the member loads are issued with the
// points' own loads.
template <int PRE, bool XP, bool MB, bool ZF>
__global__ void __launch_bounds__(kRedThreads)
    k_mx_update_gx(float *__restrict__ r, const float *__restrict__ w,
                   const uint16_t *__restrict__ dinv, const uint8_t *__restrict__ flag,
                   float *__restrict__ x, const float *__restrict__ p, int nelt,
                   double *__restrict__ part, int nblk, const CgScalars *__restrict__ s,
                   GxView G)
{
    __shared__ int32_t tabs[kGxMaxEb * kGxStride];
    const int blk = blockIdx.x;
    if (s->done) return;
    const int64_t e0 = int64_t(blk) * G.eb;
    const int ne = int(nelt - e0 < G.eb ? nelt - e0 : G.eb);
    gx_rows<kRedThreads>(tabs, G, e0, ne);
    __shared__ float vsum[kGxMaxEb * 8];
    gx_vertices<float, kRedThreads>(vsum, w, tabs, ne);
    const int np = ne * G.n3;
    const int64_t beg = e0 * G.n3;
    const float alpha = float(s->alpha);
    double a0 = 0.0, a1 = 0.0;
    constexpr int U = kGxU;
    for (int j0 = threadIdx.x; j0 < np; j0 += U * kRedThreads) {
        float rv[U], wv[U], dv[U], xv[U], pv[U], ps[U];
        uint8_t fv[U];
        bool sh[U], gm[U];
        float gv[U][4];
        int gc[U], go[U], gs[U], ge[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * kRedThreads;
            sh[u] = false;
            if (j < np) {
                const int64_t i = beg + j;
                rv[u] = r[i];
                wv[u] = w[i];
                if constexpr (MB) {
                    const uint32_t h = dinv[i];
                    dv[u] = __uint_as_float(h << 16);
                    fv[u] = uint8_t(1u << (h & 3u)) << kMultShift;
                } else {
                    fv[u] = flag[i];
                    if constexpr (PRE == NEKBONE_PRECOND_JACOBI) dv[u] = axk::dinv_at(dinv, i);
                }
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
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (sh[u]) {
                float v = gx_finish(gv[u], gc[u], go[u], gs[u], vsum + 8 * ge[u],
                                    wv[u]) +
                          0.1f * ps[u];
                if (gm[u]) v = 0.0f;
                wv[u] = v;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * kRedThreads;
            if (j >= np) continue;
            const int64_t i = beg + j;
            if constexpr (XP) x[i] = __builtin_fmaf(alpha, pv[u], xv[u]);
            if constexpr (ZF) {
                const float zi = __builtin_fmaf(-alpha, dv[u] * wv[u], rv[u]);
                r[i] = zi;
                const double rd = double(zi / dv[u]);
                const double rc = rd * inv_mult(fv[u]);
                a0 = a0 + rc * rd;
                a1 = a1 + rc * double(zi);
            } else {
                const float ri = __builtin_fmaf(-alpha, wv[u], rv[u]);
                r[i] = ri;
                const double rd = double(ri);
                const double rc = rd * inv_mult(fv[u]);
                a0 = a0 + rc * rd;
                if constexpr (PRE == NEKBONE_PRECOND_JACOBI) a1 = a1 + rc * double(ri * dv[u]);
            }
        }
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[blk] = a0;
        part[nblk + blk] = a1;
    }
}

// Reliable update, first step: apply a pending x32 += alpha p32 (the
// element kernel defers it by one iteration), then x64 += x32, x32 = 0.
__global__ void k_mx_xflush(double *__restrict__ x64, float *__restrict__ x32,
                            const float *__restrict__ p32, int64_t n,
                            const CgScalars *__restrict__ s)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float xi = x32[i];
    if (s->xpending) xi = __builtin_fmaf(float(s->alpha), p32[i], xi);
    x64[i] = x64[i] + double(xi);
    x32[i] = 0.0f;
}

// Reliable update, residual: r64 = mask(f - (A_L x64 assembled + 0.1 x64))
// (the FP64 operator of src/cg.f:96-125), r32 = r64, parts of
// rtr = (r c r) and rtz = (r c M^-1 r), with the inner iterations' own
// (bfloat16) Jacobi diagonal so that beta stays consistent.
template <int PRE>
__global__ void __launch_bounds__(kRedThreads)
    k_mx_resid(double *__restrict__ r64, float *__restrict__ r32, const double *__restrict__ f,
               const double *__restrict__ w, const double *__restrict__ x,
               const uint16_t *__restrict__ dinv, const uint8_t *__restrict__ flag, int mbit,
               int64_t n, double *__restrict__ part, int nblk)
{
    double a0 = 0.0, a1 = 0.0;
    constexpr int U = 4;  // points per thread with their loads in flight together
    const int64_t stride = int64_t(gridDim.x) * kRedThreads * U;
    for (int64_t i0 = int64_t(blockIdx.x) * kRedThreads * U + threadIdx.x; i0 < n; i0 += stride) {
        double fv[U], wv[U], xv[U];
        float dv[U];
        uint8_t fl[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + int64_t(u) * kRedThreads;
            if (i < n) {
                fv[u] = f[i];
                wv[u] = w[i];
                xv[u] = x[i];
                fl[u] = flag[i];
                if constexpr (PRE == NEKBONE_PRECOND_JACOBI) dv[u] = axk::dinv_at(dinv, i);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + int64_t(u) * kRedThreads;
            if (i < n) {
                double v = fv[u] - (wv[u] + 0.1 * xv[u]);
                if (fl[u] & mbit) v = 0.0;
                r64[i] = v;
                // Jacobi: the inner iterations carry z = d r
                if constexpr (PRE == NEKBONE_PRECOND_JACOBI)
                    r32[i] = float(v * double(dv[u]));
                else
                    r32[i] = float(v);
                const double rc = v * inv_mult(fl[u]);
                a0 = a0 + rc * v;
                if constexpr (PRE == NEKBONE_PRECOND_JACOBI) a1 = a1 + rc * (v * double(dv[u]));
            }
        }
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a0;
        part[nblk + blockIdx.x] = a1;
    }
}

// The same residual in the gather form: w = A_L x at every copy (the
// element kernel's output), summed over a shared point's copies here
// (entity table rows in LDS, the member order of the gather-form update),
// so no gather-scatter pass runs.  Runs when mx_want holds (device-driven
// refreshes).  (Folding 0.1 c x into the element kernel's w instead, so
// that x is not read here, was measured slower: 786 against 778 us for the
// pair at 32^3 N=9.)
template <int PRE>
__global__ void __launch_bounds__(kRedThreads)
    k_mx_resid_gx(double *__restrict__ r64, float *__restrict__ r32, const double *__restrict__ f,
                  const double *__restrict__ w, const double *__restrict__ x,
                  const uint16_t *__restrict__ dinv, const uint8_t *__restrict__ flag, int mbit,
                  int nelt, double *__restrict__ part, int nblk, const CgScalars *__restrict__ s,
                  GxView G)
{
    __shared__ int32_t tabs[kGxMaxEb * kGxStride];
    __shared__ double vsum[kGxMaxEb * 8];
    if (!mx_want(s)) return;
    const int blk = blockIdx.x;
    const int64_t e0 = int64_t(blk) * G.eb;
    const int ne = int(nelt - e0 < G.eb ? nelt - e0 : G.eb);
    gx_rows<kRedThreads>(tabs, G, e0, ne);
    gx_vertices<double, kRedThreads>(vsum, w, tabs, ne);
    const int np = ne * G.n3;
    const int64_t beg = e0 * G.n3;
    double a0 = 0.0, a1 = 0.0;
    constexpr int U = kGxU;
    for (int j0 = threadIdx.x; j0 < np; j0 += U * kRedThreads) {
        double fv[U], wv[U], xv[U], gv[U][4];
        float dv[U];
        uint8_t fl[U];
        bool sh[U], gm[U];
        int gc[U], go[U], gs[U], ge[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * kRedThreads;
            sh[u] = false;
            if (j < np) {
                const int64_t i = beg + j;
                fv[u] = f[i];
                wv[u] = w[i];
                xv[u] = x[i];
                fl[u] = flag[i];
                if constexpr (PRE == NEKBONE_PRECOND_JACOBI) dv[u] = axk::dinv_at(dinv, i);
                const int el = gx_divf(j, G.n3, G.inv_n3f);
                ge[u] = el;
                sh[u] = gx_issue(gv[u], gc[u], go[u], gs[u], gm[u], w, tabs + el * kGxStride, G,
                                 j - el * G.n3, i);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * kRedThreads;
            if (j >= np) continue;
            const int64_t i = beg + j;
            const double wa =
                sh[u] ? gx_finish(gv[u], gc[u], go[u], gs[u], vsum + 8 * ge[u], wv[u]) : wv[u];
            double v = fv[u] - (wa + 0.1 * xv[u]);
            if (fl[u] & mbit) v = 0.0;
            r64[i] = v;
            if constexpr (PRE == NEKBONE_PRECOND_JACOBI)
                r32[i] = float(v * double(dv[u]));
            else
                r32[i] = float(v);
            const double rc = v * inv_mult(fl[u]);
            a0 = a0 + rc * v;
            if constexpr (PRE == NEKBONE_PRECOND_JACOBI) a1 = a1 + rc * (v * double(dv[u]));
        }
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[blk] = a0;
        part[nblk + blk] = a1;
    }
}

// GATE: device-driven refreshes (runs when mx_want holds; records whether
// the FP64 residual met the tolerance and counts the refresh)
template <int PRE, bool GATE = false>
__global__ void __launch_bounds__(kFinalThreads)
    k_mx_resid_final(const double *__restrict__ part, int nblk, CgScalars *__restrict__ s)
{
    if constexpr (GATE) {
        if (!mx_want(s)) return;
    }
    double rtr = slice_sum<kFinalThreads>(part, nblk);
    double rtz = slice_sum<kFinalThreads>(part + nblk, nblk);
    block_sum2<kFinalThreads>(rtr, rtz);
    if (threadIdx.x == 0) {
        if (PRE != NEKBONE_PRECOND_JACOBI) rtz = rtr;
        s->rtr = rtr;
        s->rtr_upd = rtr;
        s->rtz1 = rtz;
        s->beta = (s->iter == 0) ? 0.0 : rtz / s->rtz2;
        s->xpending = 0;
        s->done = (rtr <= s->tol2 * s->rtr0) ? 1 : 0;
        if (!isfinite(rtr)) s->status = 1;
        if constexpr (GATE) {
            s->fin = s->done;
            s->nupd += 1;
        }
    }
}

int upd_blocks(int64_t n)
{
    const int64_t b = (n + int64_t(kRedThreads) * 16 - 1) / (int64_t(kRedThreads) * 16);
    return int(std::max<int64_t>(1, std::min<int64_t>(b, 8192)));
}

int ax32_blocks(const nekbone_problem *P) { return (P->nelt + P->mx->axv.epb - 1) / P->mx->axv.epb; }

// FP32 operator data, rebuilt when the FP64 originals changed.
int ensure_mixed(nekbone_problem *P, int precond)
{
    if (!P->mx) {
        auto *M = new MixedState();
        P->mx = M;
        const int nx = P->nx1;
        if (int rc = dalloc(&M->g, 6 * P->n)) return rc;
        if (int rc = dalloc(&M->Drow, int64_t(nx) * nx)) return rc;
        if (int rc = dalloc(&M->Dcol, int64_t(nx) * nx)) return rc;
        for (float **v : {&M->x, &M->r, &M->p, &M->w})
            if (int rc = dalloc(v, P->n)) return rc;
        std::vector<float> dr(size_t(nx) * nx), dc(size_t(nx) * nx);
        for (int i = 0; i < nx; ++i)
            for (int m = 0; m < nx; ++m) {
                dr[size_t(i * nx + m)] = float(P->dxm1[i + nx * m]);  // D(i,m)
                dc[size_t(i * nx + m)] = float(P->dxm1[m + nx * i]);  // D(m,i)
            }
        if (int rc = h2d(M->Drow, dr.data(), sizeof(float) * dr.size())) return rc;
        if (int rc = h2d(M->Dcol, dc.data(), sizeof(float) * dc.size())) return rc;
        // default FP32 shape: the FP64 default's elements per workgroup
        // with the geometric factors prefetched at entry (measured fastest
        // at N=9 on MI355X; FP32 leaves the registers for it), where that
        // shape is compiled
        M->axv = ax_default_variant(nx);
        M->axv.xu = 1;
        M->axv.pf = 1;
        if (!lookup32(nx, M->axv.epb, 1, M->axv.pf, 1)) M->axv.pf = 0;
        if (!lookup32(nx, M->axv.epb, 1, M->axv.pf, 1)) {
            M->axv.epb = axk::epb_for(nx);
            M->axv.pf = 0;
        }
    }
    MixedState *M = P->mx;
    if (M->g16 && M->g16_version != P->geom_version) {
        if (!M->gh)
            if (int rc = dalloc(&M->gh, 6 * P->n)) return rc;
        if (!M->gsc)
            if (int rc = dalloc(&M->gsc, 6 * int64_t(P->nelt))) return rc;
        const int n3 = P->nx1 * P->nx1 * P->nx1;
        hipLaunchKernelGGL(k_to_g16, dim3(6 * P->nelt), dim3(256), 0, stream(),
                           reinterpret_cast<__half *>(M->gh), M->gsc,
                           static_cast<const double *>(P->d_g), n3, int64_t(P->nelt));
        NB_HIP(hipGetLastError());
        M->g16_version = P->geom_version;
    }
    if (M->geom_version != P->geom_version) {
        hipLaunchKernelGGL(k_g_to_f32_points, dim3(grid(6 * P->n)), dim3(kThreads), 0, stream(),
                           M->g, P->d_g, P->n);
        NB_HIP(hipGetLastError());
        M->geom_version = P->geom_version;
    }
    if (precond == NEKBONE_PRECOND_JACOBI && M->dinv_version != P->dinv_version) {
        if (!M->dinv)
            if (int rc = dalloc(&M->dinv, P->n)) return rc;
        // multiplicities 1, 2, 4, 8 only (no group outside the pair, quad and
        // octet classes): the diagonal carries them
        M->mult_in_dinv = P->gs.nr == 0 ? 1 : 0;
        hipLaunchKernelGGL(k_to_bf16, dim3(grid(P->n)), dim3(kThreads), 0, stream(), M->dinv,
                           P->d_dinv, M->mult_in_dinv ? P->d_flag : nullptr, P->n);
        NB_HIP(hipGetLastError());
        M->dinv_version = P->dinv_version;
    }
    return NEKBONE_OK;
}

int launch_ax32(nekbone_problem *P, int mode, int mbit, float *w, const float *u, int gx = 0)
{
    MixedState *M = P->mx;
    AxFn32 fn = lookup32(P->nx1, M->axv.epb, mode, M->axv.pf, M->axv.xu, M->g16);
    if (!fn) return fail(NEKBONE_ERR_UNSUPPORTED, "mixed: no FP32 element kernel for this order");
    AxArgsT<float> A{};
    A.w = w;
    A.u = u;
    A.p = M->p;
    A.r = M->r;
    A.x = M->x;
    A.g = M->g;
    A.gcs = P->n;
    A.Drow = M->Drow;
    A.Dcol = M->Dcol;
    A.flag = P->d_flag;
    A.iref = P->d_iref;
    A.mbit = mbit;
    A.scal = P->d_scal;
    A.part = P->d_part;
    A.nelt = P->nelt;
    A.gh = reinterpret_cast<const __half *>(M->gh);
    A.gsc = M->gsc;
    A.gx = gx;
    const int threads = axk::round64(P->nx1 * P->nx1 * M->axv.epb);
    if (int rc = timed_begin(P, 0)) return rc;
    hipLaunchKernelGGL(fn, dim3(ax32_blocks(P)), dim3(threads), 0, stream(), A);
    NB_HIP(hipGetLastError());
    return timed_end(P, 0);
}

// Inner iteration in the gather form (cg_gather): element kernel with pap
// parts over all points, pap/alpha, the update kernel summing the shared
// points' copies, rtr/beta.  (The x-ring of the FP64 CG was measured
// neutral here: with FP16 factors at 32^3 N=9 the element kernel without
// the x update is 29 us faster and the ring's flush costs 29 us per
// iteration.)
int iteration32_gather(nekbone_problem *P, int precond, int mbit)
{
    MixedState *M = P->mx;
    const bool jac = precond == NEKBONE_PRECOND_JACOBI;
    const int64_t na = ax32_blocks(P), ng = gs_blocks(P);
    const GxView G = gx_view(P, kGxPts32);
    const int nu = int((P->nelt + G.eb - 1) / G.eb);
    double *part_upd = P->d_part + na + ng;
    if (int rc = launch_ax32(P, 1, mbit, M->w, nullptr, 1)) return rc;
    if (int rc = launch_pap_final(P, P->d_part, na)) return rc;
    if (int rc = timed_begin(P, 3)) return rc;
    const bool xp = !M->axv.xu;
#define NB_MXG2(PRE, XP, MB, ZF)                                                                \
    hipLaunchKernelGGL((k_mx_update_gx<PRE, XP, MB, ZF>), dim3(nu), dim3(kRedThreads), 0,        \
                       stream(), M->r, M->w, M->dinv, P->d_flag, M->x,                          \
                       static_cast<const float *>(M->p), P->nelt, part_upd, nu, P->d_scal, G)
#define NB_MXG(PRE, MB, ZF)                 \
    do {                                    \
        if (xp)                             \
            NB_MXG2(PRE, true, MB, ZF);     \
        else                                \
            NB_MXG2(PRE, false, MB, ZF);    \
    } while (0)
    if (jac && M->mult_in_dinv)
        NB_MXG(NEKBONE_PRECOND_JACOBI, true, true);
    else if (jac)
        NB_MXG(NEKBONE_PRECOND_JACOBI, false, true);
    else
        NB_MXG(NEKBONE_PRECOND_IDENTITY, false, false);
#undef NB_MXG
#undef NB_MXG2
    NB_HIP(hipGetLastError());
    if (int rc = timed_end(P, 3)) return rc;
    return launch_update_final(P, part_upd, nu, precond, M->axv.xu);
}

int iteration32(nekbone_problem *P, int precond, int mbit)
{
    if (cg_gather(P)) return iteration32_gather(P, precond, mbit);
    MixedState *M = P->mx;
    const bool jac = precond == NEKBONE_PRECOND_JACOBI;
    const int64_t na = ax32_blocks(P), ng = gs_blocks(P);
    const int nu = upd_blocks(P->n);
    double *part_upd = P->d_part + na + ng;
    // Jacobi runs in z form: the element kernel's identity mode reads z
    if (int rc = launch_ax32(P, 1, mbit, M->w, nullptr)) return rc;
    const bool fold = ng * kRedThreads >= na;  // as in the FP64 iteration
    if (int rc = launch_gs_cg32(P, M->w, M->p, mbit, P->d_part + na, fold ? P->d_part : nullptr,
                                fold ? na : 0))
        return rc;
    if (int rc = launch_pap_final(P, fold ? P->d_part + na : P->d_part, fold ? ng : na + ng))
        return rc;
    if (int rc = timed_begin(P, 3)) return rc;
    const bool xp = !M->axv.xu;
#define NB_MXU(PRE, XP, MB, ZF)                                                               \
    hipLaunchKernelGGL((k_mx_update<PRE, XP, MB, ZF>), dim3(nu), dim3(kRedThreads), 0, stream(), \
                       M->r, M->w, M->dinv, P->d_flag, M->x, M->p, P->n, part_upd, nu,          \
                       P->d_scal)
    if (jac && M->mult_in_dinv) {
        if (xp)
            NB_MXU(NEKBONE_PRECOND_JACOBI, true, true, true);
        else
            NB_MXU(NEKBONE_PRECOND_JACOBI, false, true, true);
    } else if (jac) {
        if (xp)
            NB_MXU(NEKBONE_PRECOND_JACOBI, true, false, true);
        else
            NB_MXU(NEKBONE_PRECOND_JACOBI, false, false, true);
    } else {
        if (xp)
            NB_MXU(NEKBONE_PRECOND_IDENTITY, true, false, false);
        else
            NB_MXU(NEKBONE_PRECOND_IDENTITY, false, false, false);
    }
#undef NB_MXU
    NB_HIP(hipGetLastError());
    if (int rc = timed_end(P, 3)) return rc;
    return launch_update_final(P, part_upd, nu, precond, M->axv.xu);
}

// x64 += x32; r64 = mask(f - A x64) in FP64; the recurrence restarts from
// it (p32 kept).
int reliable_update(nekbone_problem *P, int precond, int mbit)
{
    MixedState *M = P->mx;
    int rc = M->refresh ? launch_ax_flush(P, M->x, M->p) : NEKBONE_ERR_UNSUPPORTED;
    if (rc == NEKBONE_ERR_UNSUPPORTED) {
        hipLaunchKernelGGL(k_mx_xflush, dim3(grid(P->n)), dim3(kThreads), 0, stream(), P->d_x,
                           M->x, M->p, P->n, static_cast<const CgScalars *>(P->d_scal));
        NB_HIP(hipGetLastError());
        rc = launch_ax_local(P, P->d_w, P->d_x);
    }
    if (rc) return rc;
    if (int rc = launch_gs(P, P->d_w, nullptr)) return rc;
    const int nb = upd_blocks(P->n);
    if (precond == NEKBONE_PRECOND_JACOBI) {
        hipLaunchKernelGGL(k_mx_resid<NEKBONE_PRECOND_JACOBI>, dim3(nb), dim3(kRedThreads), 0,
                           stream(), P->d_r, M->r, P->d_f, P->d_w, P->d_x,
                           static_cast<const uint16_t *>(M->dinv), P->d_flag,
                           mbit, P->n, P->d_part, nb);
        hipLaunchKernelGGL(k_mx_resid_final<NEKBONE_PRECOND_JACOBI>, dim3(1),
                           dim3(kFinalThreads), 0, stream(), P->d_part, nb, P->d_scal);
    } else {
        hipLaunchKernelGGL(k_mx_resid<NEKBONE_PRECOND_IDENTITY>, dim3(nb), dim3(kRedThreads), 0,
                           stream(), P->d_r, M->r, P->d_f, P->d_w, P->d_x,
                           static_cast<const uint16_t *>(nullptr), P->d_flag, mbit, P->n,
                           P->d_part, nb);
        hipLaunchKernelGGL(k_mx_resid_final<NEKBONE_PRECOND_IDENTITY>, dim3(1),
                           dim3(kFinalThreads), 0, stream(), P->d_part, nb, P->d_scal);
    }
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

// Gather form: the refresh queued after every chunk of inner iterations,
// each kernel running only when mx_want holds on the device: x flush and
// w = A_L x (one element kernel), then the residual with the copies summed
// (no gather-scatter pass).
int refresh_gated(nekbone_problem *P, int precond, int mbit)
{
    MixedState *M = P->mx;
    if (int rc = launch_ax_flush(P, M->x, M->p, 1)) return rc;
    const GxView G = gx_view(P, kGxPts64);
    const int nu = int((P->nelt + G.eb - 1) / G.eb);
    if (precond == NEKBONE_PRECOND_JACOBI) {
        hipLaunchKernelGGL(k_mx_resid_gx<NEKBONE_PRECOND_JACOBI>, dim3(nu), dim3(kRedThreads), 0,
                           stream(), P->d_r, M->r, P->d_f, P->d_w,
                           static_cast<const double *>(P->d_x),
                           static_cast<const uint16_t *>(M->dinv), P->d_flag, mbit, P->nelt,
                           P->d_part, nu, P->d_scal, G);
        hipLaunchKernelGGL((k_mx_resid_final<NEKBONE_PRECOND_JACOBI, true>), dim3(1),
                           dim3(kFinalThreads), 0, stream(), P->d_part, nu, P->d_scal);
    } else {
        hipLaunchKernelGGL(k_mx_resid_gx<NEKBONE_PRECOND_IDENTITY>, dim3(nu), dim3(kRedThreads),
                           0, stream(), P->d_r, M->r, P->d_f, P->d_w,
                           static_cast<const double *>(P->d_x),
                           static_cast<const uint16_t *>(nullptr), P->d_flag, mbit, P->nelt,
                           P->d_part, nu, P->d_scal, G);
        hipLaunchKernelGGL((k_mx_resid_final<NEKBONE_PRECOND_IDENTITY, true>), dim3(1),
                           dim3(kFinalThreads), 0, stream(), P->d_part, nu, P->d_scal);
    }
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int read_scalars(CgScalars &s, const nekbone_problem *P)
{
    NB_HIP(hipMemcpyAsync(&s, P->d_scal, sizeof s, hipMemcpyDeviceToHost, stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

}  // namespace

void free_mixed(nekbone_problem *P)
{
    if (!P->mx) return;
    MixedState *M = P->mx;
    for (float **v : {&M->g, &M->Drow, &M->Dcol, &M->x, &M->r, &M->p, &M->w}) dfree(*v);
    dfree(M->dinv);
    dfree(M->gh);
    dfree(M->gsc);
    delete M;
    P->mx = nullptr;
}

}  // namespace nb

using namespace nb;

extern "C" int nekbone_cg_mixed_d(nekbone_problem *P, double *x, const double *f, double *r,
                                  double tol, int maxiter, int precond, double delta, int *iters,
                                  int *updates, double *final_rtr)
{
    if (int rc = check(P)) return rc;
    if (!(tol > 0.0)) return fail(NEKBONE_ERR_ARG, "cg_mixed: tol must be > 0");
    if (maxiter < 0) return fail(NEKBONE_ERR_ARG, "cg_mixed: maxiter < 0");
    // default residual reduction between FP64 updates: 1e-3 for the FP32
    // operator; 5e-3 (FP16 factors) and 3e-2 (FP16 matrix cores) for the
    // FP16 ones, whose ~5e-4 relative operator error otherwise costs the
    // inner iteration its convergence rate (measured at 32^3 N=9 deformed
    // to 1e-10: 94 inner iterations with these, against 111 and 427 with
    // 1e-3; 5e-3 needs 4 FP64 updates, 1e-2 five)
    if (delta == 0.0)
        delta = !P->mx ? 1e-3
                : P->mx->g16 == NEKBONE_MIXED_G16 ? 5e-3
                : P->mx->g16 == NEKBONE_MIXED_F16_MFMA ? 3e-2
                : 1e-3;
    if (!(delta > 0.0 && delta < 1.0)) return fail(NEKBONE_ERR_ARG, "cg_mixed: need 0 < delta < 1");
    if (precond != NEKBONE_PRECOND_JACOBI && precond != NEKBONE_PRECOND_IDENTITY)
        return fail(NEKBONE_ERR_UNSUPPORTED, "cg_mixed: preconditioner not available");
    if (precond == NEKBONE_PRECOND_JACOBI && !P->jacobi)
        return fail(NEKBONE_ERR_STATE, "cg_mixed: Jacobi preconditioner not set up");
    if (!P->c_is_mult)
        return fail(NEKBONE_ERR_UNSUPPORTED,
                    "cg_mixed: needs c = 1/multiplicity (the inner kernels derive c from it)");
    if (int rc = ensure_mixed(P, precond)) return rc;
    if (int rc = launch_to_dev(P, P->d_f, f)) return rc;
    const int mbit0 = next_mask_bit(P);
    if (int rc = launch_cg_init(P, P->d_x, P->d_r, P->d_f, mbit0, precond, tol * tol, false,
                                false))
        return rc;
    const int mbit = next_mask_bit(P);
    MixedState *M = P->mx;
    const double delta2 = delta * delta;
    hipLaunchKernelGGL(k_mx_start, dim3(grid(P->n)), dim3(kThreads), 0, stream(), M->r, M->x,
                       M->p, P->d_r, P->n, P->d_scal,
                       precond == NEKBONE_PRECOND_JACOBI ? static_cast<const uint16_t *>(M->dinv)
                                                         : nullptr,
                       delta2, maxiter);
    NB_HIP(hipGetLastError());

    const int chunk = 8;
    int nupd = 0;
    CgScalars s{};
    if (cg_gather(P) && M->refresh && P->arith != NEKBONE_ARITH_FMA) {
        // Device-driven refreshes: every chunk of inner iterations is
        // followed by the gated refresh and a copy of the scalars into
        // pinned memory, and the next chunk is queued before the host waits
        // for that copy, so the GPU never waits for the host.  Inner
        // iterations after the recurrence has converged return at once
        // until the refresh has checked the FP64 residual; if that leaves
        // fewer than maxiter iterations done when nothing is queued, more
        // chunks follow.
        int queued = 0, posted = 0, seen = 0;
        auto queue = [&]() -> int {
            const int todo = std::min(chunk, maxiter - queued);
            for (int k = 0; k < todo; ++k)
                if (int rc = iteration32(P, precond, mbit)) return rc;
            queued += todo;
            if (int rc = refresh_gated(P, precond, mbit)) return rc;
            return scal_post(P, posted++ & 1);
        };
        if (int rc = queue()) return rc;
        for (;;) {
            if (queued < maxiter)
                if (int rc = queue()) return rc;
            if (int rc = scal_wait(seen++ & 1, s)) return rc;
            if (s.fin || s.status) break;
            if (seen == posted) {
                if (s.iter >= maxiter) break;
                queued = s.iter;
                if (int rc = queue()) return rc;
            }
        }
        nupd = s.nupd;
    } else {
        if (int rc = read_scalars(s, P)) return rc;
        while (!s.done && s.iter < maxiter && !s.status) {
            const int todo = std::min(chunk, maxiter - s.iter);
            for (int k = 0; k < todo; ++k)
                if (int rc = iteration32(P, precond, mbit)) return rc;
            if (int rc = read_scalars(s, P)) return rc;
            if (s.status) break;
            if (s.done || s.iter >= maxiter || s.rtr <= delta2 * s.rtr_upd) {
                if (int rc = reliable_update(P, precond, mbit)) return rc;
                ++nupd;
                if (int rc = read_scalars(s, P)) return rc;
            }
        }
    }
    if (x)
        if (int rc = launch_to_ref(P, x, P->d_x)) return rc;
    if (r)
        if (int rc = launch_to_ref(P, r, P->d_r)) return rc;
    NB_HIP(hipStreamSynchronize(stream()));
    if (iters) *iters = s.iter;
    if (updates) *updates = nupd;
    if (final_rtr) *final_rtr = s.rtr;
    if (s.status) return fail(NEKBONE_ERR_NUMERIC, "cg_mixed: non-positive or non-finite pap");
    if (!s.done)
        return fail(NEKBONE_NOT_CONVERGED,
                    "cg_mixed: stopped at maxiter with the FP64 rtr > tol^2 rtr0");
    return NEKBONE_OK;
}

extern "C" int nekbone_set_mixed_refresh(nekbone_problem *P, int fused)
{
    if (int rc = check(P)) return rc;
    if (fused != 0 && fused != 1) return fail(NEKBONE_ERR_ARG, "mixed: refresh must be 0 or 1");
    if (int rc = ensure_mixed(P, NEKBONE_PRECOND_IDENTITY)) return rc;
    P->mx->refresh = fused;
    return NEKBONE_OK;
}

extern "C" int nekbone_set_mixed_inner(nekbone_problem *P, int mode)
{
    if (int rc = check(P)) return rc;
    if (mode < NEKBONE_MIXED_F32 || mode > NEKBONE_MIXED_F16_MFMA)
        return fail(NEKBONE_ERR_ARG, "mixed: unknown inner operator");
    if (int rc = ensure_mixed(P, NEKBONE_PRECOND_IDENTITY)) return rc;
    MixedState *M = P->mx;
    if (mode != M->g16) {
        // the first shape compiled for this operator: the default one, else
        // one element per workgroup (the only shape of the matrix-core kernel)
        AxVariant v{axk::epb_for(P->nx1), 0, 1};
        if (!lookup32(P->nx1, v.epb, 1, v.pf, v.xu, mode)) v = AxVariant{1, 0, 1};
        if (!lookup32(P->nx1, v.epb, 1, v.pf, v.xu, mode) ||
            !lookup32(P->nx1, v.epb, 0, v.pf, 0, mode))
            return fail(NEKBONE_ERR_UNSUPPORTED, "mixed: inner operator not built for this order");
        M->axv = v;
    }
    M->g16 = mode;
    return ensure_mixed(P, NEKBONE_PRECOND_IDENTITY);
}

extern "C" int nekbone_set_mixed_g16(nekbone_problem *P, int on)
{
    return nekbone_set_mixed_inner(P, on ? NEKBONE_MIXED_G16 : NEKBONE_MIXED_F32);
}

extern "C" int nekbone_get_variant_f32(nekbone_problem *P, int *epb, int *pf, int *xu)
{
    if (int rc = check(P)) return rc;
    if (int rc = ensure_mixed(P, NEKBONE_PRECOND_IDENTITY)) return rc;
    if (epb) *epb = P->mx->axv.epb;
    if (pf) *pf = P->mx->axv.pf;
    if (xu) *xu = P->mx->axv.xu;
    return NEKBONE_OK;
}

extern "C" int nekbone_set_variant_f32(nekbone_problem *P, int epb, int pf, int xu)
{
    if (int rc = check(P)) return rc;
    if (int rc = ensure_mixed(P, NEKBONE_PRECOND_IDENTITY)) return rc;
    if (!lookup32(P->nx1, epb, 1, pf, xu ? 1 : 0, P->mx->g16) ||
        !lookup32(P->nx1, epb, 0, pf, 0, P->mx->g16))
        return fail(NEKBONE_ERR_UNSUPPORTED, "set_variant_f32: shape not compiled for this order");
    P->mx->axv = AxVariant{epb, pf, xu ? 1 : 0};
    return NEKBONE_OK;
}

extern "C" int nekbone_autotune_mixed(nekbone_problem *P, int precond, double *best_ms)
{
    if (int rc = check(P)) return rc;
    if (precond == NEKBONE_PRECOND_JACOBI && !P->jacobi)
        return fail(NEKBONE_ERR_STATE, "autotune_mixed: Jacobi not set up");
    if (precond != NEKBONE_PRECOND_JACOBI && precond != NEKBONE_PRECOND_IDENTITY)
        return fail(NEKBONE_ERR_ARG, "autotune_mixed: identity or Jacobi");
    if (!P->c_is_mult) return fail(NEKBONE_ERR_UNSUPPORTED, "autotune_mixed: needs c = 1/mult");
    if (int rc = ensure_mixed(P, precond)) return rc;
    MixedState *M = P->mx;
    std::vector<AxVariant> cands;
    for (const auto &kv : table32()) {
        int nx, epb, mode, pf, xu, g16;
        std::tie(nx, epb, mode, pf, xu, g16) = kv.first;
        if (nx == P->nx1 && mode == 1 && g16 == M->g16) cands.push_back(AxVariant{epb, pf, xu});
    }
    const AxVariant keep = M->axv;
    const int mbit = P->mask_mode == NEKBONE_MASK_DIRICHLET ? 1 : 2;
    const int warm = 2, reps = 6;
    if (int rc = tune_rhs(P, 12345)) return rc;
    hipEvent_t e0, e1;
    NB_HIP(hipEventCreate(&e0));
    NB_HIP(hipEventCreate(&e1));
    double best = 1e300;
    AxVariant win = keep;
    int rc = NEKBONE_OK;
    for (const AxVariant &v : cands) {
        M->axv = v;
        rc = launch_cg_init(P, P->d_x, P->d_r, P->d_f, 1, precond, 0.0, false, false);
        if (!rc) {
            hipLaunchKernelGGL(k_mx_start, dim3(grid(P->n)), dim3(kThreads), 0, stream(), M->r,
                               M->x, M->p, P->d_r, P->n, P->d_scal,
                               precond == NEKBONE_PRECOND_JACOBI
                                   ? static_cast<const uint16_t *>(M->dinv)
                                   : nullptr,
                               0.0, 0);
            if (hipGetLastError() != hipSuccess) rc = NEKBONE_ERR_HIP;
        }
        for (int i = 0; i < warm && !rc; ++i) rc = iteration32(P, precond, mbit);
        if (!rc && hipEventRecord(e0, stream()) != hipSuccess) rc = NEKBONE_ERR_HIP;
        for (int i = 0; i < reps && !rc; ++i) rc = iteration32(P, precond, mbit);
        if (!rc && hipEventRecord(e1, stream()) != hipSuccess) rc = NEKBONE_ERR_HIP;
        if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = NEKBONE_ERR_HIP;
        float ms = 0.f;
        if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = NEKBONE_ERR_HIP;
        if (rc) break;
        if (ms / reps < best) {
            best = ms / reps;
            win = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        M->axv = keep;
        return rc == NEKBONE_ERR_HIP ? fail(rc, "autotune_mixed: HIP error") : rc;
    }
    M->axv = win;
    if (best_ms) *best_ms = best;
    return NEKBONE_OK;
}

extern "C" int nekbone_ax_local_f32(nekbone_problem *P, float *w, const float *u)
{
    if (int rc = check(P)) return rc;
    if (!w || !u) return fail(NEKBONE_ERR_ARG, "ax_local_f32: null array");
    if (int rc = ensure_mixed(P, NEKBONE_PRECOND_IDENTITY)) return rc;
    MixedState *M = P->mx;
    const int64_t n = P->n;
    const int n3 = P->nx1 * P->nx1 * P->nx1;
    // stage through the FP32 work vectors (their CG contents are not kept)
    NB_HIP(hipMemcpyAsync(M->w, u, sizeof(float) * size_t(n), hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL(k_perm32<true>, dim3(grid(n)), dim3(kThreads), 0, stream(), M->r, M->w,
                       P->d_perm, n3, n);
    NB_HIP(hipGetLastError());
    if (int rc = launch_ax32(P, 0, 0, M->p, M->r)) return rc;
    hipLaunchKernelGGL(k_perm32<false>, dim3(grid(n)), dim3(kThreads), 0, stream(), M->w, M->p,
                       P->d_perm, n3, n);
    NB_HIP(hipGetLastError());
    NB_HIP(hipMemcpyAsync(w, M->w, sizeof(float) * size_t(n), hipMemcpyDeviceToHost, stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}
