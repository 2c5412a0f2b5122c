// Element operator w_e = A_L u_e = D^T G D u_e for every element
// (ax_e, src/cg.f:142-212), FP64, for gfx950.
//
// Design (one workgroup = EPB elements, NX*NX threads per element):
//  * Every 1-D contraction is done by one thread along a whole grid line.
//    All lanes of a wave then use the same coefficient D(i,m) at the same
//    time, so D is read through the scalar cache (SGPR operands) and the
//    only LDS traffic is the element data itself: about 20 LDS accesses
//    per point instead of ~120 for a thread-per-point scheme.
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
//  * Geometric factors are stored element-blocked SoA, g[e][c][p], so the
//    six loads of a t-line step are each one coalesced row.
//  * PMODE fuses the CG direction update p = beta*p + z (add2s1,
//    src/math.f:534-541) into the load of u: 1 = identity (z = r),
//    2 = Jacobi (z = r*dinv).  p is written back once.
#include "nb_internal.h"

namespace nb {
namespace {

template <bool FMA>
__device__ __forceinline__ double madd(double acc, double a, double b)
{
    if constexpr (FMA)
        return __builtin_fma(a, b, acc);
    else
        return acc + a * b;
}

constexpr int round64(int x) { return (x + 63) / 64 * 64; }

template <int NX, int EPB, bool FMA, int PMODE>
__global__ void __launch_bounds__(round64(NX *NX *EPB))
    k_ax_local(double *__restrict__ w, const double *__restrict__ u, double *__restrict__ p,
               const double *__restrict__ r, const double *__restrict__ dinv,
               const double *__restrict__ g, const double *__restrict__ D,
               const CgScalars *__restrict__ scal, int nelt)
{
    constexpr int N2 = NX * NX, N3 = NX * NX * NX;
    __shared__ double sA[EPB * N3];
    __shared__ double sB[EPB * N3];

    const int tid = threadIdx.x;
    const int el = tid / N2;
    const int ab = tid - el * N2;
    const int a = ab % NX, b = ab / NX;
    const int64_t e = int64_t(blockIdx.x) * EPB + el;
    const bool act = (el < EPB) && (e < nelt);
    double *A = sA + (act ? el : 0) * N3;
    double *B = sB + (act ? el : 0) * N3;
    const int64_t base = e * N3;
    if constexpr (PMODE != 0) {
        if (scal->done) return;  // converged tolerance run: no more work
    }

    // D(i,m) = D[i + NX*m]; indices are compile-time constants after
    // unrolling and the address is uniform, so these become scalar loads.
    auto Dm = [&](int i, int m) { return D[i + NX * m]; };

    double tt[NX];  // ut, later the t-part of the transposed operator
    if (act) {
        double uc[NX];
        if constexpr (PMODE == 0) {
#pragma unroll
            for (int k = 0; k < NX; ++k) uc[k] = u[base + a + NX * b + N2 * k];
        } else {
            const double beta = scal->beta;
#pragma unroll
            for (int k = 0; k < NX; ++k) {
                const int64_t q = base + a + NX * b + N2 * k;
                double z = r[q];
                if constexpr (PMODE == 2) z = z * dinv[q];
                double pn;
                if constexpr (FMA)
                    pn = __builtin_fma(beta, p[q], z);
                else
                    pn = beta * p[q] + z;
                p[q] = pn;
                uc[k] = pn;
            }
        }
#pragma unroll
        for (int k = 0; k < NX; ++k) A[a + NX * b + N2 * k] = uc[k];
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            double s = uc[0] * Dm(k, 0);
#pragma unroll
            for (int m = 1; m < NX; ++m) s = madd<FMA>(s, uc[m], Dm(k, m));
            tt[k] = s;
        }
    }
    __syncthreads();

    double lr[NX], ls[NX];
    if (act) {
#pragma unroll
        for (int m = 0; m < NX; ++m) lr[m] = A[m + NX * a + N2 * b];  // r-line (j,k)=(a,b)
#pragma unroll
        for (int m = 0; m < NX; ++m) ls[m] = A[a + NX * m + N2 * b];  // s-line (i,k)=(a,b)
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int i = 0; i < NX; ++i) {
            double s = Dm(i, 0) * lr[0];
#pragma unroll
            for (int m = 1; m < NX; ++m) s = madd<FMA>(s, Dm(i, m), lr[m]);
            B[i + NX * a + N2 * b] = s;  // ur
        }
#pragma unroll
        for (int j = 0; j < NX; ++j) {
            double s = ls[0] * Dm(j, 0);
#pragma unroll
            for (int m = 1; m < NX; ++m) s = madd<FMA>(s, ls[m], Dm(j, m));
            A[a + NX * j + N2 * b] = s;  // us
        }
    }
    __syncthreads();

    if (act) {
        const double *ge = g + e * 6 * N3;
        double wt[NX];
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const int q = a + NX * b + N2 * k;
            const double ur = B[q], us = A[q], ut = tt[k];
            const double g1 = ge[q], g2 = ge[N3 + q], g3 = ge[2 * N3 + q];
            const double g4 = ge[3 * N3 + q], g5 = ge[4 * N3 + q], g6 = ge[5 * N3 + q];
            double wr, ws;
            if constexpr (FMA) {
                wr = __builtin_fma(g3, ut, __builtin_fma(g2, us, g1 * ur));
                ws = __builtin_fma(g5, ut, __builtin_fma(g4, us, g2 * ur));
                wt[k] = __builtin_fma(g6, ut, __builtin_fma(g5, us, g3 * ur));
            } else {
                wr = g1 * ur + g2 * us + g3 * ut;
                ws = g2 * ur + g4 * us + g5 * ut;
                wt[k] = g3 * ur + g5 * us + g6 * ut;
            }
            B[q] = wr;
            A[q] = ws;
        }
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            double s = wt[0] * Dm(0, k);
#pragma unroll
            for (int m = 1; m < NX; ++m) s = madd<FMA>(s, wt[m], Dm(m, k));
            tt[k] = s;
        }
    }
    __syncthreads();

    if (act) {
        // r-line and s-line are each read and rewritten by their owner
        // only, so no barrier is needed between the reads and the writes.
#pragma unroll
        for (int m = 0; m < NX; ++m) lr[m] = B[m + NX * a + N2 * b];
#pragma unroll
        for (int i = 0; i < NX; ++i) {
            double s = Dm(0, i) * lr[0];
#pragma unroll
            for (int m = 1; m < NX; ++m) s = madd<FMA>(s, Dm(m, i), lr[m]);
            B[i + NX * a + N2 * b] = s;  // Tr
        }
#pragma unroll
        for (int m = 0; m < NX; ++m) ls[m] = A[a + NX * m + N2 * b];
#pragma unroll
        for (int j = 0; j < NX; ++j) {
            double s = ls[0] * Dm(0, j);
#pragma unroll
            for (int m = 1; m < NX; ++m) s = madd<FMA>(s, ls[m], Dm(m, j));
            A[a + NX * j + N2 * b] = s;  // Ts
        }
    }
    __syncthreads();

    if (act) {
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const int q = a + NX * b + N2 * k;
            w[base + q] = (B[q] + A[q]) + tt[k];
        }
    }
}

// Elements per workgroup for each NX: enough threads for >= 2 waves and
// LDS well inside 160 KB.
constexpr int epb_for(int nx)
{
    return nx <= 3 ? 16 : nx <= 4 ? 8 : nx <= 5 ? 4 : nx <= 8 ? 2 : 1;
}

using AxFn = void (*)(double *, const double *, double *, const double *, const double *,
                      const double *, const double *, const CgScalars *, int);

template <int NX>
AxFn pick(bool fma, int pmode)
{
    constexpr int E = epb_for(NX);
    if (fma) {
        if (pmode == 0) return k_ax_local<NX, E, true, 0>;
        if (pmode == 1) return k_ax_local<NX, E, true, 1>;
        return k_ax_local<NX, E, true, 2>;
    }
    if (pmode == 0) return k_ax_local<NX, E, false, 0>;
    if (pmode == 1) return k_ax_local<NX, E, false, 1>;
    return k_ax_local<NX, E, false, 2>;
}

AxFn select(int nx, bool fma, int pmode, int *epb)
{
    switch (nx) {
#define NB_CASE(N)            \
    case N:                   \
        *epb = epb_for(N);    \
        return pick<N>(fma, pmode);
        NB_CASE(2) NB_CASE(3) NB_CASE(4) NB_CASE(5) NB_CASE(6) NB_CASE(7) NB_CASE(8)
        NB_CASE(9) NB_CASE(10) NB_CASE(11) NB_CASE(12) NB_CASE(13) NB_CASE(14)
        NB_CASE(15) NB_CASE(16)
#undef NB_CASE
    default:
        return nullptr;
    }
}

}  // namespace

int launch_ax_local(nekbone_problem *P, double *w, const double *u, int pmode, double *p,
                    const double *r)
{
    int epb = 1;
    AxFn fn = select(P->nx1, P->arith == NEKBONE_ARITH_FMA, pmode, &epb);
    if (!fn) return fail(NEKBONE_ERR_UNSUPPORTED, "ax: nx1 outside 2..16");
    if (pmode == 2 && !P->d_dinv) return fail(NEKBONE_ERR_STATE, "ax: Jacobi not set up");
    const int threads = round64(P->nx1 * P->nx1 * epb);
    const int blocks = (P->nelt + epb - 1) / epb;
    if (int rc = timed_begin(P, 0)) return rc;
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(threads), 0, stream(), w, u, p, r, P->d_dinv,
                       P->d_g, P->d_D, P->d_scal, P->nelt);
    NB_HIP(hipGetLastError());
    return timed_end(P, 0);
}

}  // namespace nb
