// Element operator kernel template (included by the nb_ax*.hip files).
// w_e = A_L u_e = D^T G D u_e for every element
// (ax_e, src/cg.f:142-212), FP64, for gfx950, plus the fused CG variant.
//
// Design (one workgroup = EPB elements, NX*NX threads per element):
//  * Every 1-D contraction is done by one thread along a whole grid line.
//    All lanes of a wave use the same coefficient at the same time, so D
//    comes through the scalar cache as SGPR operands and the only LDS
//    traffic is the element data: about 20 LDS accesses per point instead
//    of ~120 for a thread-per-point scheme.
//  * D is stored twice, row-major (Drow[i*NX+m] = D(i,m)) and
//    column-major (Dcol[i*NX+m] = D(m,i)), so every contraction reads one
//    contiguous row of NX doubles through the constant address space
//    (scalar loads).  An empty asm statement on the uniform row pointer
//    keeps the compiler from hoisting all NX^2 coefficients into SGPRs at
//    once, which would spill them into VGPR lanes and double the VALU work.
//  * A thread owns line (a,b) in three roles: t-line (i,j)=(a,b), r-line
//    (j,k)=(a,b), s-line (i,k)=(a,b).  The t-direction never leaves
//    registers; r and s lines are exchanged through two LDS buffers of
//    NX^3 doubles per element (16 KB at NX=10).
//  * Every sum runs over m = 0..NX-1 in order starting from the first
//    product, i.e. the order of mxfN (src/mxm_std.f:68-631):
//      ur(i,j,k) = sum_m D(i,m) u(m,j,k)      mxm(D,m1,u,m1,ur,m2)
//      us(i,j,k) = sum_m u(i,m,k) D(j,m)      mxm(u(,,k),m1,Dt,m1,us,m1)
//      ut(i,j,k) = sum_m u(i,j,m) D(k,m)      mxm(u,m2,Dt,m1,ut,m1)
//      w = (sum_m D(m,i) wr(m,j,k) + sum_m ws(i,m,k) D(m,j))
//          + sum_m wt(i,j,m) D(m,k)           local_grad3_t
//    and wr = (g1 ur + g2 us) + g3 ut etc. (src/cg.f:156-163).  With
//    FMA=false no contraction happens (the library is compiled with
//    -ffp-contract=off), so results are bit-identical to the reference.
//  * Geometric factors are stored as six mesh-wide planes, g[c][e][p]
//    (factor c of point p of element e at c*n + e*N3 + p): the six loads
//    of a t-line step are each one coalesced row, and the whole launch
//    reads seven long sequential streams (u and the six planes), which
//    HBM serves faster than the same bytes blocked per element.  PF=1
//    issues the loads at kernel entry (deeper memory pipelining, more
//    VGPRs); PF=0 issues them in the geometric phase (higher occupancy);
//    PF=2 is PF=0 compiled for 5 waves per SIMD in FP64 (at most 96
//    VGPRs, the most that 16 KB of LDS per element allows at one element
//    per workgroup) and for 8 in FP32 (at most 64 VGPRs).  PF=3 loads
//    them one t-step ahead of their use (software pipelining).
//
// CG mode (MODE 1) fuses, per point of the thread's t-line:
//   prologue  [XU] x += alpha_prev p_prev     (add2s2, src/cg.f:63, one
//                                              iteration late; with the
//                                              x-ring, pout set, x is left
//                                              to the update kernel)
//             p = beta p + z                  (add2s1, src/cg.f:56; z is
//                                              the vector the CG carries:
//                                              r with the identity, M^-1 r
//                                              with Jacobi)
//   epilogue  points not shared with another element are final after
//             the element operator: w = w + 0.1 p (src/cg.f:118), mask
//             (src/cg.f:119), and their part of pap = sum (w c) p
//             (src/cg.f:59) with c = 1 there.  Shared points are left
//             for the gather-scatter kernel, which finishes them, or in
//             the gather form (gx) for the vector-update kernel, which
//             sums their copies itself; then this kernel's pap parts
//             cover the shared points as well.
// The variant (EPB, PF, XU) is chosen per order by the autotuner.
#pragma once

#include <map>
#include <tuple>
#include <vector>

#include <hip/hip_fp16.h>

#include "nb_cg_device.h"

namespace nb {
namespace axk {

// Fused multiply-add in the operand type (__builtin_fma alone would
// promote float operands to double).
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c)
{
    return __builtin_fma(a, b, c);
}

template <bool FMA, typename T>
__device__ __forceinline__ T madd(T acc, T a, T b)
{
    if constexpr (FMA)
        return fma_t(a, b, acc);
    else
        return acc + a * b;
}

constexpr int round64(int x) { return (x + 63) / 64 * 64; }

// sum_m row[m] * x[m], m ascending, first term a plain product.
template <int NX, bool FMA, typename T>
__device__ __forceinline__ T dot_row(const T *__restrict__ base, int row, const T (&x)[NX])
{
    typedef const __attribute__((address_space(4))) T *const_ptr;
    const_ptr p = (const_ptr)(base + row * NX);
    asm volatile("" : "+s"(p));
    T s = p[0] * x[0];
#pragma unroll
    for (int m = 1; m < NX; ++m) s = madd<FMA>(s, p[m], x[m]);
    return s;
}

// ---- FP16 matrix-core contractions (MF variant, FP32 storage) ---------
// v_mfma_f32_16x16x16_f16, one wave: lane l holds A[l%16][4(l/16)+v],
// B[4(l/16)+v][l%16] and C[4(l/16)+v][l%16], v = 0..3 (checked on the
// device against a host product).
typedef _Float16 mf_half4 __attribute__((ext_vector_type(4)));
typedef float mf_float4 __attribute__((ext_vector_type(4)));

// B fragment of the 16x16 matrix Bm[m][c] = D(c,m) (TR = false, the
// gradient) or D(m,c) (TR = true, its transpose), zero outside NX x NX.
// Drow[i*NX+m] = D(i,m).
template <int NX, bool TR>
__device__ __forceinline__ mf_half4 mf_dfrag(const float *__restrict__ Drow, int lane)
{
    mf_half4 b;
    const int c = lane & 15;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int m = 4 * (lane >> 4) + v;
        float d = 0.f;
        if (m < NX && c < NX) d = TR ? Drow[m * NX + c] : Drow[c * NX + m];
        b[v] = (_Float16)d;
    }
    return b;
}

// One 16-row tile of a contraction along the direction of stride SIG
// (1: r, NX: s, NX^2: t) of an element held as FP16 in src (reference
// point order): rows are the NX^2 lines of that direction, the line of
// row rho starting at line_base(rho); out[point] = scale * sum_m
// src[line + SIG m] Bm[m][c], written (ACC = false) or added (ACC =
// true) at point line + SIG c.
template <int NX, int SIG>
__device__ __forceinline__ int mf_line_base(int rho)
{
    if constexpr (SIG == 1) return NX * rho;           // rho = j + NX k
    else if constexpr (SIG == NX) return (rho % NX) + NX * NX * (rho / NX);  // rho = i + NX k
    else return rho;                                    // rho = i + NX j
}

template <int NX, int SIG, bool ACC>
__device__ __forceinline__ void mf_tile(float *__restrict__ out, const _Float16 *__restrict__ src,
                                        mf_half4 bfrag, int tile, int lane, float scale)
{
    constexpr int N2 = NX * NX;
    mf_half4 a;
    const int ra = 16 * tile + (lane & 15);
    const int la = ra < N2 ? mf_line_base<NX, SIG>(ra) : 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int m = 4 * (lane >> 4) + v;
        a[v] = (ra < N2 && m < NX) ? src[la + SIG * m] : (_Float16)0.f;
    }
    mf_float4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x16f16(a, bfrag, c, 0, 0, 0);
    const int col = lane & 15;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int rc = 16 * tile + 4 * (lane >> 4) + v;
        if (rc < N2 && col < NX) {
            const int q = mf_line_base<NX, SIG>(rc) + SIG * col;
            if constexpr (ACC)
                out[q] = out[q] + scale * c[v];
            else
                out[q] = scale * c[v];
        }
    }
}

// Largest |v| over the workgroup (every thread returns it): wave
// butterflies of max, then the wave maxima through LDS.  One barrier
// for the exchange, one so that the helper can run again.
template <int NT>
__device__ __forceinline__ float mf_block_max(float v)
{
    constexpr int NW = NT / 64;
    __shared__ float sm[NW];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    float m = sm[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) m = fmaxf(m, sm[w]);
    __syncthreads();
    return m;
}

// Jacobi diagonal of the FP32 inner iterations of the mixed solver:
// bfloat16 bits (the FP32 exponent range, 8-bit significand), only ever
// used as a preconditioner.
__device__ __forceinline__ float dinv_at(const uint16_t *d, int64_t i)
{
    return __uint_as_float(uint32_t(d[i]) << 16);
}

// T is the storage and arithmetic type of the operator (double; float
// for the inner solve of the mixed-precision CG).  Dot-product partials
// and the CG scalars are FP64 for both.
template <typename T>
struct AxArgsT {
    T *w;
    const T *u;     // MODE 0 input
    T *p;           // CG: direction, updated in place (read only when pout is set)
    T *pout;        // CG, x-ring: where the new direction goes (null: in place)
    const T *r;     // CG: z (r with the identity, M^-1 r with Jacobi)
    T *x;           // CG: solution, updated one iteration late (XU)
    const T *g;     // FP64: planes g[c * gcs + e * N3 + p]; FP32: g[6 * (e * N3 + p) + c]
    const T *Drow, *Dcol;
    const uint8_t *flag;
    const int16_t *iref;  // reference position of each device position (N3)
    const __half *gh;     // G16: geometric factors in FP16 (six per point),
                          // g = gh * gsc[c * nelt + e]
    const float *gsc;     // G16: per factor and element scale
    int mbit;
    const CgScalars *scal;
    double *part;        // CG: per-block pap partials
    int nelt;
    int64_t gcs;         // distance between factor planes (local points n)
    int rev;             // walk the elements downward (see sweep_block)
    int gx;              // CG, gather form: the pap parts cover shared points too (below)
    int gate;            // MODE 3: run only when mx_want holds (device-driven refreshes)
    // MODE 3 (first pass of the FP64 residual refresh of the mixed-
    // precision CG): x = x + the FP32 increment x32 (+ alpha p32 when
    // pending), x32 = 0, then w = A_L x at every point as in MODE 0.
    float *x32;
    const float *p32;
};
using AxArgs = AxArgsT<double>;

// Body of the element kernel for workgroup blk of one problem (shared by
// the single-problem kernel, the batched ones and the work-queue kernel).
// sA, sB: two LDS planes of EPB*NX^3 values; NTB: threads of the
// workgroup (0: round64(NX^2*EPB), the kernel's own shape; more when a
// work-queue kernel runs several orders with one block size; the extra
// threads own no line).
template <typename T, int NX, int EPB, bool FMA, int MODE, int PF, bool XU, bool G16 = false,
          bool MF = false, int NTB = 0>
__device__ __forceinline__ void ax_core(const AxArgsT<T> &A_, const int blk, T *__restrict__ sA,
                                        T *__restrict__ sB)
{
    constexpr int N2 = NX * NX, N3 = NX * NX * NX;
    constexpr int NT = NTB ? NTB : round64(N2 * EPB);
    static_assert(NT >= N2 * EPB && NT % 64 == 0, "workgroup too small for its elements");
    static_assert(!MF || (EPB == 1 && G16 && sizeof(T) == 4 && NX <= 16 && NTB == 0),
                  "FP16 matrix-core variant: FP32 storage, FP16 factors, one element per group");
    __shared__ T sC[MF ? N3 : 1];          // MF: us, then the accumulated w
    __shared__ _Float16 sH[MF ? N3 : 1];   // MF: FP16 input, then FP16 wt

    const int tid = threadIdx.x;
    const int el = tid / N2;
    const int ab = tid - el * N2;
    const int a = ab % NX, b = ab / NX;
    const int64_t e = int64_t(blk) * EPB + el;
    const bool act = (el < EPB) && (e < A_.nelt);
    T *A = sA + (act ? el : 0) * N3;
    T *B = sB + (act ? el : 0) * N3;
    const int64_t base = e * N3;
    const T *__restrict__ Drow = A_.Drow;
    const T *__restrict__ Dcol = A_.Dcol;
    // Factor layouts: FP64, six mesh-wide planes (factor c of point q at
    // ge[c*gcs + q]); FP32 and FP16 (inner solves of the mixed CG), the six
    // factors of a point contiguous (g[6*(e*N3 + q) + c]), read with one
    // 8-byte load per two factors (FP32) or one 4-byte load (FP16), which
    // the narrow types need to stream near the HBM rate.
    constexpr bool AOS = sizeof(T) == 4;
    const int64_t gcs = A_.gcs;
    const T *ge = A_.g + (AOS ? 6 : 1) * e * N3;
    // G16 (FP32 inner solves only): factor c at point q is gh*scale
    const __half *gh = G16 ? A_.gh + 6 * e * N3 : nullptr;
    T gsc[6] = {T(1), T(1), T(1), T(1), T(1), T(1)};
    if constexpr (G16) {
        if (act) {
#pragma unroll
            for (int c = 0; c < 6; ++c) gsc[c] = T(A_.gsc[c * int64_t(A_.nelt) + e]);
        }
    }
    // the six factors of point q of the element
    auto gload6 = [&](int q, T(&g)[6]) {
        if constexpr (G16) {
            const uint32_t *p = reinterpret_cast<const uint32_t *>(gh + 6 * q);
#pragma unroll
            for (int h = 0; h < 3; ++h) {
                const uint32_t v = p[h];
                g[2 * h] = T(__half2float(__ushort_as_half(uint16_t(v & 0xffffu)))) * gsc[2 * h];
                g[2 * h + 1] = T(__half2float(__ushort_as_half(uint16_t(v >> 16)))) * gsc[2 * h + 1];
            }
        } else if constexpr (AOS) {
            const float2 *p = reinterpret_cast<const float2 *>(ge + 6 * q);
#pragma unroll
            for (int h = 0; h < 3; ++h) {
                const float2 v = p[h];
                g[2 * h] = v.x;
                g[2 * h + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) g[c] = ge[c * gcs + q];
        }
    };

    if constexpr (MODE == 1) {
        if (A_.scal->done) return;  // converged tolerance run: nothing to do
    }
    if constexpr (MODE == 3) {
        if (A_.gate && !mx_want(A_.scal)) return;  // no refresh due
    }

    constexpr bool PRE = PF == 1;
    T gv[PRE ? NX : 1][6];
    if constexpr (PRE) {
        if (act) {
#pragma unroll
            for (int k = 0; k < NX; ++k) gload6(a + NX * b + N2 * k, gv[k]);
        }
    }

    // Linear phase: the element's N3 points in device order are split as
    // d = ab + N2*m, m = 0..NX-1 (N3 = NX*N2), so every global access of
    // the per-point vectors is one contiguous run per wave whatever the
    // device order is.  ref[m] is the reference position of point d (a
    // 2-byte table shared by all elements, so it stays in cache).
    int ref[NX];
    if (act) {
#pragma unroll
        for (int m = 0; m < NX; ++m) ref[m] = A_.iref[ab + N2 * m];
    }
    // the points' flags, loaded with the vectors (four to a register) so
    // that the epilogue does not wait for a fresh memory round trip
    // measured: FP64 gains, FP32 loses
    constexpr bool FLPRE = MODE == 1 && sizeof(T) == 8;
    uint32_t flp[FLPRE ? (NX + 3) / 4 : 1] = {};
    if constexpr (FLPRE) {
        if (act) {
#pragma unroll
            for (int m = 0; m < NX; ++m)
                flp[m / 4] |= uint32_t(A_.flag[base + ab + N2 * m]) << (8 * (m % 4));
        }
    }

    T pd[MODE == 1 ? NX : 1];  // CG: the new direction at the linear points
    if (act) {
        if constexpr (MODE == 0) {
#pragma unroll
            for (int m = 0; m < NX; ++m) A[ref[m]] = A_.u[base + ab + N2 * m];
        } else if constexpr (MODE == 3) {
            const bool xp = A_.scal->xpending != 0;
            const float al = float(A_.scal->alpha);
#pragma unroll
            for (int m = 0; m < NX; ++m) {
                const int64_t q = base + ab + N2 * m;
                float xi = A_.x32[q];
                if (xp) xi = __builtin_fmaf(al, A_.p32[q], xi);
                const T xv = A_.x[q] + T(xi);
                A_.x[q] = xv;
                A_.x32[q] = 0.0f;
                A[ref[m]] = xv;
            }
        } else {
            const T beta = T(A_.scal->beta);
            // x-ring: the first iteration has no previous direction (its
            // ring slot may hold anything), beta = 0 there
            const bool pnone = A_.pout && A_.scal->iter == 0;
            T po[NX], z[NX];
#pragma unroll
            for (int m = 0; m < NX; ++m) {
                const int64_t q = base + ab + N2 * m;
                po[m] = pnone ? T(0) : A_.p[q];
                z[m] = A_.r[q];
            }
            if constexpr (XU) {
                if (A_.scal->xpending) {  // the previous iteration's alpha p
                    const T alpha = T(A_.scal->alpha);
#pragma unroll
                    for (int m = 0; m < NX; ++m) {
                        const int64_t q = base + ab + N2 * m;
                        if constexpr (FMA)
                            A_.x[q] = fma_t(alpha, po[m], A_.x[q]);
                        else
                            A_.x[q] = A_.x[q] + alpha * po[m];
                    }
                }
            }
#pragma unroll
            for (int m = 0; m < NX; ++m) {
                T pn;
                if constexpr (FMA)
                    pn = fma_t(beta, po[m], z[m]);
                else
                    pn = beta * po[m] + z[m];
                (A_.pout ? A_.pout : A_.p)[base + ab + N2 * m] = pn;
                pd[m] = pn;
                A[ref[m]] = pn;
            }
        }
    }
    __syncthreads();

    if constexpr (MF) {
        // FP16 matrix-core operator (no counterpart in the reference; the
        // inner operator of the mixed-precision CG).  The element's input,
        // held in A (FP32), is scaled by its largest magnitude and rounded
        // to FP16; the three gradient contractions run on MFMA with FP32
        // accumulation; the geometric product is FP32; the three
        // transposed contractions take the products scaled by their
        // largest magnitude, rounded to FP16, and accumulate w in FP32.
        const int lane = tid & 63, wv = tid >> 6;
        constexpr int NW = NT / 64, TPC = (N2 + 15) / 16;  // waves, tiles per contraction
        _Float16 *H = sH;  // FP16 input
        float umax = 0.f;
        T uc[NX];
        if (act) {
#pragma unroll
            for (int k = 0; k < NX; ++k) {
                uc[k] = A[a + NX * b + N2 * k];
                umax = fmaxf(umax, fabsf(uc[k]));
            }
        }
        const float su = mf_block_max<NT>(umax);
        const float isu = su > 0.f ? 1.f / su : 1.f;
        if (act) {
#pragma unroll
            for (int k = 0; k < NX; ++k) H[a + NX * b + N2 * k] = (_Float16)(uc[k] * isu);
        }
        __syncthreads();
        const mf_half4 bg = mf_dfrag<NX, false>(A_.Drow, lane);
        for (int t = wv; t < 3 * TPC; t += NW) {
            if (t < TPC)
                mf_tile<NX, 1, false>(sA, H, bg, t, lane, su);             // ur -> A
            else if (t < 2 * TPC)
                mf_tile<NX, NX, false>(sC, H, bg, t - TPC, lane, su);      // us -> C
            else
                mf_tile<NX, NX * NX, false>(sB, H, bg, t - 2 * TPC, lane, su);  // ut -> B
        }
        __syncthreads();
        T wr[NX], ws[NX], wt[NX];
        float wmax = 0.f;
        if (act) {
#pragma unroll
            for (int k = 0; k < NX; ++k) {
                const int q = a + NX * b + N2 * k;
                const T ur = sA[q], us = sC[q], ut = sB[q];
                T gq[6];
                gload6(q, gq);
                const T g1 = gq[0], g2 = gq[1], g3 = gq[2], g4 = gq[3], g5 = gq[4], g6 = gq[5];
                wr[k] = fma_t(g3, ut, fma_t(g2, us, g1 * ur));
                ws[k] = fma_t(g5, ut, fma_t(g4, us, g2 * ur));
                wt[k] = fma_t(g6, ut, fma_t(g5, us, g3 * ur));
                wmax = fmaxf(wmax, fmaxf(fabsf(wr[k]), fmaxf(fabsf(ws[k]), fabsf(wt[k]))));
            }
        }
        const float sw = mf_block_max<NT>(wmax);  // its barrier also ends every read above
        const float isw = sw > 0.f ? 1.f / sw : 1.f;
        _Float16 *HR = reinterpret_cast<_Float16 *>(sA);
        _Float16 *HS = HR + N3;  // second half of A's bytes
        _Float16 *HT = sH;
        if (act) {
#pragma unroll
            for (int k = 0; k < NX; ++k) {
                const int q = a + NX * b + N2 * k;
                HR[q] = (_Float16)(wr[k] * isw);
                HS[q] = (_Float16)(ws[k] * isw);
                HT[q] = (_Float16)(wt[k] * isw);
            }
        }
        __syncthreads();
        const mf_half4 bt = mf_dfrag<NX, true>(A_.Drow, lane);
        for (int t = wv; t < TPC; t += NW) mf_tile<NX, 1, false>(sC, HR, bt, t, lane, sw);
        __syncthreads();
        for (int t = wv; t < TPC; t += NW) mf_tile<NX, NX, true>(sC, HS, bt, t, lane, sw);
        __syncthreads();
        for (int t = wv; t < TPC; t += NW) mf_tile<NX, NX * NX, true>(sC, HT, bt, t, lane, sw);
        __syncthreads();
    } else {
        T uc[NX];  // the thread's t-line of the input
        if (act) {
#pragma unroll
            for (int k = 0; k < NX; ++k) uc[k] = A[a + NX * b + N2 * k];
        }

        T lr[NX], ls[NX];
        // PF=3: the factors are loaded one t-step ahead of their use, the
        // first step's before the r and s contractions, so a load round trip
        // overlaps a step's arithmetic instead of stalling it (the empty asm
        // that keeps D in scalar registers is a scheduling barrier, so the
        // compiler would not hoist them itself)
        constexpr bool PIPE = PF == 3;
        T gnx[PIPE ? 6 : 1];
        if constexpr (PIPE) {
            if (act) gload6(a + NX * b, gnx);
        }
        if (act) {
#pragma unroll
            for (int m = 0; m < NX; ++m) lr[m] = A[m + NX * a + N2 * b];  // r-line (j,k)=(a,b)
#pragma unroll
            for (int m = 0; m < NX; ++m) ls[m] = A[a + NX * m + N2 * b];  // s-line (i,k)=(a,b)
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int i = 0; i < NX; ++i) B[i + NX * a + N2 * b] = dot_row<NX, FMA>(Drow, i, lr);
#pragma unroll
            for (int j = 0; j < NX; ++j) A[a + NX * j + N2 * b] = dot_row<NX, FMA>(Drow, j, ls);
        }
        __syncthreads();

        T wt[NX];
        if (act) {
#pragma unroll
            for (int k = 0; k < NX; ++k) {
                const int q = a + NX * b + N2 * k;
                const T ur = B[q], us = A[q];
                T gq[6];
                if constexpr (PIPE) {
#pragma unroll
                    for (int c = 0; c < 6; ++c) gq[c] = gnx[c];
                    if (k + 1 < NX) gload6(q + N2, gnx);
                }
                const T ut = dot_row<NX, FMA>(Drow, k, uc);
                if constexpr (PRE) {
#pragma unroll
                    for (int c = 0; c < 6; ++c) gq[c] = gv[k][c];
                } else if constexpr (!PIPE) {
                    gload6(q, gq);
                }
                const T g1 = gq[0], g2 = gq[1], g3 = gq[2], g4 = gq[3], g5 = gq[4], g6 = gq[5];
                T wr, ws;
                if constexpr (FMA) {
                    wr = fma_t(g3, ut, fma_t(g2, us, g1 * ur));
                    ws = fma_t(g5, ut, fma_t(g4, us, g2 * ur));
                    wt[k] = fma_t(g6, ut, fma_t(g5, us, g3 * ur));
                } else {
                    wr = g1 * ur + g2 * us + g3 * ut;
                    ws = g2 * ur + g4 * us + g5 * ut;
                    wt[k] = g3 * ur + g5 * us + g6 * ut;
                }
                B[q] = wr;
                A[q] = ws;
            }
        }
        __syncthreads();

        if (act) {
            // r-line and s-line are each read and rewritten by their owner
            // only, so no barrier is needed between the reads and the writes.
#pragma unroll
            for (int m = 0; m < NX; ++m) lr[m] = B[m + NX * a + N2 * b];
#pragma unroll
            for (int i = 0; i < NX; ++i) B[i + NX * a + N2 * b] = dot_row<NX, FMA>(Dcol, i, lr);
#pragma unroll
            for (int m = 0; m < NX; ++m) ls[m] = A[a + NX * m + N2 * b];
#pragma unroll
            for (int j = 0; j < NX; ++j) A[a + NX * j + N2 * b] = dot_row<NX, FMA>(Dcol, j, ls);
        }
        __syncthreads();

        // w at the thread's t-line points, kept in place of the consumed
        // partials (each entry is read and rewritten by its owner only)
        if (act) {
#pragma unroll
            for (int k = 0; k < NX; ++k) {
                const int q = a + NX * b + N2 * k;
                A[q] = (B[q] + A[q]) + dot_row<NX, FMA>(Dcol, k, wt);
            }
        }
        __syncthreads();
    }
    double acc = 0.0;
    if (act) {
        // positions read again (cache hits) rather than kept live through
        // the contractions
        const int16_t *iref = A_.iref;
        asm volatile("" : "+s"(iref));
#pragma unroll
        for (int m = 0; m < NX; ++m) {
            const int64_t q = base + ab + N2 * m;
            const T wl = MF ? sC[iref[ab + N2 * m]] : A[iref[ab + N2 * m]];
            if constexpr (MODE == 0 || MODE == 3) {
                A_.w[q] = wl;
            } else {
                const uint8_t fl = FLPRE ? uint8_t(flp[m / 4] >> (8 * (m % 4))) : A_.flag[q];
                if (fl & kFlagShared) {
                    A_.w[q] = wl;  // finished by the gather-scatter or update kernel
                    if (A_.gx && !(fl & A_.mbit)) {
                        // gather form: pap = sum_i c_i p_i w_i over the
                        // copies of a shared group G equals
                        // sum_{j in G} (p_j w_L,j + 0.1 c_j p_j^2), since
                        // the copies of p agree, sum_G c = 1, and p = 0
                        // where w is masked; so every element adds its own
                        // copies' terms and no kernel after the sum needs
                        // to form pap
                        const double pq = double(pd[m]);
                        acc = acc + pq * (double(wl) + 0.1 * inv_mult(fl) * pq);
                    }
                } else {
                    T v = wl + T(0.1) * pd[m];
                    if (fl & A_.mbit) v = T(0);
                    A_.w[q] = v;
                    acc = acc + double(v) * double(pd[m]);
                }
            }
        }
    }
    if constexpr (MODE == 1) {
        // deterministic block sum of the pap partials: wave butterflies,
        // then the wave sums in wave order
        constexpr int NW = NT / 64;
        __shared__ double red[NW];
        acc = wave_sum(acc);
        if ((tid & 63) == 0) red[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            double s = red[0];
#pragma unroll
            for (int w = 1; w < NW; ++w) s = s + red[w];
            A_.part[blk] = s;
        }
    }
}

template <typename T, int NX, int EPB, bool FMA, int MODE, int PF, bool XU, bool G16 = false,
          bool MF = false>
__device__ __forceinline__ void ax_body(const AxArgsT<T> &A_, const int blk)
{
    constexpr int N3 = NX * NX * NX;
    __shared__ T sA[EPB * N3];
    __shared__ T sB[EPB * N3];
    ax_core<T, NX, EPB, FMA, MODE, PF, XU, G16, MF>(A_, blk, sA, sB);
}

template <typename T, int NX, int EPB, bool FMA, int MODE, int PF, bool XU, bool G16 = false,
          bool MF = false>
__global__ void __launch_bounds__(round64(NX *NX *EPB))
    __attribute__((amdgpu_waves_per_eu(PF == 2 ? (sizeof(T) == 4 ? 8 : 5) : 1, 8)))
    k_ax(AxArgsT<T> A_)
{
    ax_body<T, NX, EPB, FMA, MODE, PF, XU, G16, MF>(A_,
                                                   sweep_block(blockIdx.x, gridDim.x, A_.rev));
}

// Batched form: problem q owns workgroups [start[q], start[q+1]); its
// arguments are read once per workgroup.  Each problem sees exactly the
// workgroups (and partial sums) of its own single-problem launch.
template <typename T, int NX, int EPB, bool FMA, int MODE, int PF, bool XU>
__global__ void __launch_bounds__(round64(NX *NX *EPB))
    k_ax_batch(const AxArgsT<T> *__restrict__ args, const int *__restrict__ start, int nprob)
{
    const int b = blockIdx.x;
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (start[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const AxArgsT<T> A = args[lo];
    ax_body<T, NX, EPB, FMA, MODE, PF, XU>(A, b - start[lo]);
}

// Work-queue form of the batched element kernel: one launch for the
// problems of several orders.  Item i of the queue is (problem, workgroup
// of that problem's own launch); workgroup b takes items b, b + gridDim.x,
// b + 2 gridDim.x, ... (persistent), and the queue lists the highest
// orders first so that the longest items start first.  Each item runs the
// body of the problem's own kernel shape (E6..E12 elements per workgroup
// at orders 6, 8, 10, 12: the batched workload's), so every value is the
// one the problem's single solve gives with that shape.  NTB threads per
// workgroup (enough for every order's lines; the rest own none); the two
// LDS planes come from the dynamic allocation, 2 * max(E * NX^3) doubles.
// One kernel carries the register allocation of its largest body (161-233
// VGPRs uncapped against 96-128 for one order alone); capped at 128 (4
// waves per SIMD) it spills 25-33 registers and ran fastest of the caps
// measured (3 waves per SIMD and uncapped were slower).
struct QItem {
    int prob, blk;
};

template <int NTB, bool FMA, bool XU, int E6, int E8, int E10, int E12>
__global__ void __launch_bounds__(NTB) __attribute__((amdgpu_waves_per_eu(4, 8)))
    k_ax_queue(const AxArgsT<double> *__restrict__ args, const int *__restrict__ pnx,
               const QItem *__restrict__ items, int nitems)
{
    extern __shared__ double qlds[];
    for (int i = blockIdx.x; i < nitems; i += gridDim.x) {
        const QItem q = items[i];
        const AxArgsT<double> A = args[q.prob];
        switch (pnx[q.prob]) {
        case 6:
            ax_core<double, 6, E6, FMA, 1, 0, XU, false, false, NTB>(A, q.blk, qlds,
                                                                     qlds + E6 * 216);
            break;
        case 8:
            ax_core<double, 8, E8, FMA, 1, 0, XU, false, false, NTB>(A, q.blk, qlds,
                                                                     qlds + E8 * 512);
            break;
        case 10:
            ax_core<double, 10, E10, FMA, 1, 0, XU, false, false, NTB>(A, q.blk, qlds,
                                                                       qlds + E10 * 1000);
            break;
        case 12:
            ax_core<double, 12, E12, FMA, 1, 0, XU, false, false, NTB>(A, q.blk, qlds,
                                                                       qlds + E12 * 1728);
            break;
        default:
            break;
        }
        __syncthreads();  // the next item writes the LDS planes again
    }
}

using AxFnQ = void (*)(const AxArgsT<double> *, const int *, const QItem *, int);
// A compiled tile set of the work-queue kernel: threads per workgroup and
// elements per workgroup at orders 6, 8, 10, 12; fn[fma][xu].
struct QShape {
    int ntb, e[4];
    AxFnQ fn[2][2];
};
constexpr int kQOrders[4] = {6, 8, 10, 12};
// Defined in nb_ax_iq.hip.
void regq_all(std::vector<QShape> &t);

// Default elements per workgroup for each NX: >= 2 waves per group and
// LDS well inside 160 KB.
constexpr int epb_for(int nx)
{
    return nx <= 3 ? 16 : nx <= 4 ? 8 : nx <= 5 ? 4 : nx <= 8 ? 2 : 1;
}

using AxFn = void (*)(AxArgs);

// Kernel table: (nx, epb, fma, mode, pf, xu) -> function.
using Key = std::tuple<int, int, int, int, int, int>;
using Table = std::map<Key, AxFn>;

template <int NX, int EPB, int PF>
void reg_modes(Table &t)
{
    t[Key(NX, EPB, 0, 0, PF, 0)] = k_ax<double, NX, EPB, false, 0, PF, false>;
    t[Key(NX, EPB, 1, 0, PF, 0)] = k_ax<double, NX, EPB, true, 0, PF, false>;
    t[Key(NX, EPB, 0, 1, PF, 0)] = k_ax<double, NX, EPB, false, 1, PF, false>;
    t[Key(NX, EPB, 1, 1, PF, 0)] = k_ax<double, NX, EPB, true, 1, PF, false>;
    t[Key(NX, EPB, 0, 1, PF, 1)] = k_ax<double, NX, EPB, false, 1, PF, true>;
    t[Key(NX, EPB, 1, 1, PF, 1)] = k_ax<double, NX, EPB, true, 1, PF, true>;
}

// x flush + FP64 operator of the mixed-precision CG's residual refresh
// (MODE 3): the default shape, exact arithmetic.  Key (nx, epb, 0, 3, pf, 0),
// pf 0, and 3 (factors one t-step ahead) for the orders with every shape.
template <int NX>
void reg_refresh(Table &t, bool full)
{
    constexpr int E = epb_for(NX);
    t[Key(NX, E, 0, 3, 0, 0)] = k_ax<double, NX, E, false, 3, 0, false>;
    if (full) t[Key(NX, E, 0, 3, 3, 0)] = k_ax<double, NX, E, false, 3, 3, false>;
}

// Orders the benchmark configurations use get the full variant space;
// the others get the default shape.
template <int NX>
void reg_order(Table &t, bool full)
{
    constexpr int E = epb_for(NX);
    reg_refresh<NX>(t, full);
    reg_modes<NX, E, 0>(t);
    if (full) {
        reg_modes<NX, E, 1>(t);
        reg_modes<NX, E, 3>(t);
        reg_modes<NX, 2 * E, 0>(t);
        reg_modes<NX, 2 * E, 1>(t);
        if constexpr (E == 1) reg_modes<NX, E, 2>(t);
    }
}

// Batched CG kernels (double, mode 1), same keys as the table.
using AxFnB = void (*)(const AxArgs *, const int *, int);
using TableB = std::map<Key, AxFnB>;

template <int NX, int EPB, int PF>
void regb_modes(TableB &t)
{
    t[Key(NX, EPB, 0, 1, PF, 0)] = k_ax_batch<double, NX, EPB, false, 1, PF, false>;
    t[Key(NX, EPB, 0, 1, PF, 1)] = k_ax_batch<double, NX, EPB, false, 1, PF, true>;
    t[Key(NX, EPB, 1, 1, PF, 0)] = k_ax_batch<double, NX, EPB, true, 1, PF, false>;
    t[Key(NX, EPB, 1, 1, PF, 1)] = k_ax_batch<double, NX, EPB, true, 1, PF, true>;
}

template <int NX>
void regb_order(TableB &t)
{
    constexpr int E = epb_for(NX);
    regb_modes<NX, E, 0>(t);
    // the batched workload's orders also get the doubled shape (the
    // single-problem table has it too), for the batch autotuner, and the
    // default shape with the factors loaded one t-step ahead
    if constexpr (NX == 6 || NX == 8 || NX == 10 || NX == 12) {
        regb_modes<NX, 2 * E, 0>(t);
        regb_modes<NX, E, 3>(t);
    }
}

// FP32 kernels (inner iterations of the mixed-precision CG), always with
// fused multiply-adds.  Key: (nx, epb, mode, pf, xu).
using AxFn32 = void (*)(AxArgsT<float>);
// Key: (nx, epb, mode, pf, xu, g16)
using Key32 = std::tuple<int, int, int, int, int, int>;
using Table32 = std::map<Key32, AxFn32>;

template <int NX, int EPB, int PF>
void reg32_modes(Table32 &t)
{
    t[Key32(NX, EPB, 0, PF, 0, 0)] = k_ax<float, NX, EPB, true, 0, PF, false>;
    t[Key32(NX, EPB, 1, PF, 0, 0)] = k_ax<float, NX, EPB, true, 1, PF, false>;
    t[Key32(NX, EPB, 1, PF, 1, 0)] = k_ax<float, NX, EPB, true, 1, PF, true>;
    t[Key32(NX, EPB, 0, PF, 0, 1)] = k_ax<float, NX, EPB, true, 0, PF, false, true>;
    t[Key32(NX, EPB, 1, PF, 1, 1)] = k_ax<float, NX, EPB, true, 1, PF, true, true>;
}

template <int NX>
void reg32_order(Table32 &t, bool full)
{
    constexpr int E = epb_for(NX);
    reg32_modes<NX, E, 0>(t);
    if (full) {
        reg32_modes<NX, E, 1>(t);
        reg32_modes<NX, E, 3>(t);
        reg32_modes<NX, 2 * E, 0>(t);
        reg32_modes<NX, 2 * E, 1>(t);
        reg32_modes<NX, 2 * E, 3>(t);
        reg32_modes<NX, E, 2>(t);
        reg32_modes<NX, 2 * E, 2>(t);
    }
}

// FP16 matrix-core kernels of the mixed solver (key g16 = 2): one element
// per workgroup, factors in FP16; plain operator and the CG modes with
// the x update fused.
template <int NX>
void reg32_mf(Table32 &t)
{
    t[Key32(NX, 1, 0, 0, 0, 2)] = k_ax<float, NX, 1, true, 0, 0, false, true, true>;
    t[Key32(NX, 1, 1, 0, 1, 2)] = k_ax<float, NX, 1, true, 1, 0, true, true, true>;
}

// Defined in the instantiation files nb_ax_i*.hip and nb_ax32_*.hip.
void reg32_low(Table32 &t);
void reg32_high(Table32 &t);
void reg32_mf_all(Table32 &t);
void regb_all(TableB &t);
void reg_low(Table &t);
void reg_high(Table &t);
void reg_6(Table &t);
void reg_8(Table &t);
void reg_10(Table &t);
void reg_12(Table &t);

}  // namespace axk
}  // namespace nb
