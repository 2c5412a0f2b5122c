// Gather-scatter, fused CG vector kernels, deterministic reductions and
// set-up kernels.  All FP64; the library is compiled with
// -ffp-contract=off so each expression rounds exactly as written.
#include "nb_device.h"

namespace nb {
namespace {

// ---------------------------------------------------------------------
// dssum: per shared group, t = u[i0] + u[i1] + ... in ascending local
// index, then t written to every member (src/jl/gs_local.c:60-87).
// With CG=true each member is finished as well: w = t + 0.1 p
// (src/cg.f:118), mask (src/cg.f:119), and its part of
// pap = sum (w*c)*p with c = 1/(group size) (src/cg.f:59).
// One launch covers all size classes; every workgroup handles one class,
// so its lanes run the same trip count.
// ---------------------------------------------------------------------


// T = double, or float for the inner solve of the mixed-precision CG (its
// pap partials are accumulated in FP64 all the same).
template <typename T, bool CG>
__device__ __forceinline__ void gs_finish(T *__restrict__ w, const T *__restrict__ p,
                                          const uint8_t *__restrict__ flag, int mbit, int m, T t,
                                          double c, double &acc)
{
    if constexpr (CG) {
        const T pm = p[m];
        T v = t + T(0.1) * pm;
        if (flag[m] & mbit) v = T(0);
        w[m] = v;
        acc = acc + double(v) * c * double(pm);
    } else {
        w[m] = t;
    }
}

template <typename T, int S, bool CG>
__device__ __forceinline__ void gs_sum(T *__restrict__ w, const T *__restrict__ p,
                                       const uint8_t *__restrict__ flag, int mbit,
                                       const int (&m)[S], double &acc)
{
    T t = w[m[0]];
#pragma unroll
    for (int q = 1; q < S; ++q) t += w[m[q]];
    const double c = 1. / double(S);
#pragma unroll
    for (int q = 0; q < S; ++q) gs_finish<T, CG>(w, p, flag, mbit, m[q], t, c, acc);
}

template <typename T, int S, bool CG>
__device__ __forceinline__ void gs_group(T *__restrict__ w, const T *__restrict__ p,
                                         const uint8_t *__restrict__ flag, int mbit,
                                         const int32_t *__restrict__ list,
                                         const int32_t *__restrict__ runs, int R, int64_t nrun,
                                         int64_t g, double &acc)
{
    int m[S];
    gs_members<S>(list, runs, R, nrun, g, m);
    gs_sum<T, S, CG>(w, p, flag, mbit, m, acc);
}

// kPairs pairs per thread (groups g, g + kRedThreads, ... of the block's
// range): every load of all of them is issued before the first store,
// since the compiler may not move a load of a later pair above a store of
// an earlier one.  Each pair is summed and finished as gs_sum does.  Two
// measured best at 32^3 elements, N=9 (one and four: within 2%, FP64 and FP32).
constexpr int kPairs = 2;
template <typename T, bool CG>
__device__ __forceinline__ void gs_pairs(T *__restrict__ w, const T *__restrict__ p,
                                         const uint8_t *__restrict__ flag, int mbit,
                                         const GsArgs &G, int64_t g0, double &acc)
{
    int m[kPairs][2];
    bool v[kPairs];
#pragma unroll
    for (int h = 0; h < kPairs; ++h) {
        const int64_t g = g0 + int64_t(h) * kRedThreads;
        v[h] = g < G.n2;
        m[h][0] = m[h][1] = 0;
        if (v[h]) gs_members<2>(G.m2, G.m2r, G.r2, G.n2r, g, m[h]);
    }
    T a[kPairs][2], pm[kPairs][2];
    uint8_t fl[kPairs][2];
#pragma unroll
    for (int h = 0; h < kPairs; ++h) {
        if (v[h]) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                a[h][q] = w[m[h][q]];
                if constexpr (CG) {
                    pm[h][q] = p[m[h][q]];
                    fl[h][q] = flag[m[h][q]];
                }
            }
        }
    }
#pragma unroll
    for (int h = 0; h < kPairs; ++h) {
        if (!v[h]) continue;
        T t = a[h][0];
        t += a[h][1];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if constexpr (CG) {
                T vq = t + T(0.1) * pm[h][q];
                if (fl[h][q] & mbit) vq = T(0);
                w[m[h][q]] = vq;
                acc = acc + double(vq) * 0.5 * double(pm[h][q]);
            } else {
                w[m[h][q]] = t;
            }
        }
    }
}

template <typename T, bool CG>
__device__ __forceinline__ void gs_body(T *__restrict__ w, const T *__restrict__ p,
                                        const uint8_t *__restrict__ flag, int mbit,
                                        const GsArgs &G, double *__restrict__ part,
                                        const CgScalars *__restrict__ scal, const int blk0)
{
    if (scal && scal->done) return;
    int blk = blk0;
    double acc = 0.0;
    if (blk < G.b2) {
        gs_pairs<T, CG>(w, p, flag, mbit, G, int64_t(blk) * (kPairs * kRedThreads) + threadIdx.x,
                        acc);
    } else if ((blk -= G.b2) < G.b4) {
        const int64_t g = int64_t(blk) * kRedThreads + threadIdx.x;
        if (g < G.n4) gs_group<T, 4, CG>(w, p, flag, mbit, G.m4, G.m4r, G.r4, G.n4r, g, acc);
    } else if ((blk -= G.b4) < G.b8) {
        const int64_t g = int64_t(blk) * kRedThreads + threadIdx.x;
        if (g < G.n8) gs_group<T, 8, CG>(w, p, flag, mbit, G.m8, nullptr, 1, 0, g, acc);
    } else {
        blk -= G.b8;
        const int64_t g = int64_t(blk) * kRedThreads + threadIdx.x;
        if (g < G.nr) {
            const int a = G.roff[g], b = G.roff[g + 1];
            T t = w[G.ridx[a]];
            for (int q = a + 1; q < b; ++q) t += w[G.ridx[q]];
            const double c = 1. / double(b - a);
            for (int q = a; q < b; ++q) gs_finish<T, CG>(w, p, flag, mbit, G.ridx[q], t, c, acc);
        }
    }
    if constexpr (CG) {
        if (G.pre) {
            const int64_t j = int64_t(blk0) * kRedThreads + threadIdx.x;
            if (j < G.npre) acc = acc + G.pre[j];
        }
        double dummy = 0.0;
        block_sum2(acc, dummy);
        if (threadIdx.x == 0) part[blk0] = acc;
    }
}

template <typename T, bool CG>
__global__ void __launch_bounds__(kRedThreads)
    k_gs_cls(T *__restrict__ w, const T *__restrict__ p,
             const uint8_t *__restrict__ flag, int mbit, GsArgs G, double *__restrict__ part,
             const CgScalars *__restrict__ scal)
{
    gs_body<T, CG>(w, p, flag, mbit, G, part, scal, sweep_block(blockIdx.x, gridDim.x, G.rev));
}

// Batched fused CG gather-scatter: problem q owns workgroups
// [start[q], start[q+1]).
__global__ void __launch_bounds__(kRedThreads)
    k_gs_cg_batch(const GsCgArgs *__restrict__ args, const int *__restrict__ start, int nprob)
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
    const GsCgArgs A = args[lo];
    gs_body<double, true>(A.w, A.p, A.flag, A.mbit, A.G, A.part, A.scal, b - start[lo]);
}



// ---------------------------------------------------------------------
// After dssum: w = w + 0.1 p (add2s2, src/cg.f:118), then the Dirichlet
// mask (src/cg.f:119), optionally with the block partials of
// pap = sum (w*c)*p (glsc3, src/cg.f:59).
// ---------------------------------------------------------------------
template <bool PAP>
__global__ void __launch_bounds__(kRedThreads)
    k_post(double *__restrict__ w, const double *__restrict__ p, const double *__restrict__ c,
           const uint8_t *__restrict__ flag, int mbit, int64_t n, double *__restrict__ part,
           const CgScalars *__restrict__ scal)
{
    if (scal && scal->done) return;
    double acc = 0.0, dummy = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kRedThreads;
    for (int64_t i = int64_t(blockIdx.x) * kRedThreads + threadIdx.x; i < n; i += stride) {
        double v = w[i] + 0.1 * p[i];
        if (flag[i] & mbit) v = 0.0;
        w[i] = v;
        if constexpr (PAP) acc = acc + v * c[i] * p[i];
    }
    if constexpr (PAP) {
        block_sum2(acc, dummy);
        if (threadIdx.x == 0) part[blockIdx.x] = acc;
    }
}

// maskit by the per-point list flags (src/cg.f:224-229 zeroes the listed
// points; the flag bit marks exactly those points).
__global__ void k_mask_flag(double *__restrict__ w, const uint8_t *__restrict__ flag, int mbit,
                            int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && (flag[i] & mbit)) w[i] = 0.0;
}

// glsc3 block partials: sum (a*b)*m (src/math.f:603-617).
__global__ void __launch_bounds__(kRedThreads)
    k_glsc3_part(const double *__restrict__ a, const double *__restrict__ b,
                 const double *__restrict__ m, int64_t n, double *__restrict__ part)
{
    double acc = 0.0, dummy = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kRedThreads;
    for (int64_t i = int64_t(blockIdx.x) * kRedThreads + threadIdx.x; i < n; i += stride)
        acc = acc + a[i] * b[i] * m[i];
    block_sum2(acc, dummy);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Sum of nblk partials in a fixed order: thread t adds partials t,
// t + kRedThreads, ..., then the block sum.
__global__ void __launch_bounds__(kRedThreads)
    k_final_scalar(const double *__restrict__ part, int nblk, double *__restrict__ out)
{
    double a0 = 0.0, dummy = 0.0;
    for (int i = threadIdx.x; i < nblk; i += kRedThreads) a0 = a0 + part[i];
    block_sum2(a0, dummy);
    if (threadIdx.x == 0) *out = a0;
}

// ---------------------------------------------------------------------
// Set-up kernels
// ---------------------------------------------------------------------

// Brick factors g1=g4=g6=w_i w_j w_k, others 0 (src/prox_setup.f:27-52),
// stored as six planes g[c][e][p] (factor c at c*n + e*nxyz + p).
__global__ void k_brick_g(double *__restrict__ g, const double *__restrict__ wgll, int nx,
                          int64_t nelt)
{
    const int64_t nxyz = int64_t(nx) * nx * nx;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nxyz * nelt) return;
    const int64_t e = i / nxyz;
    const int p = int(i - e * nxyz);
    const int a = p % nx, b = (p / nx) % nx, k = p / (nx * nx);
    const double v = wgll[a] * wgll[b] * wgll[k];
    const int64_t n = nxyz * nelt;
    double *ge = g + e * nxyz;
    ge[p] = v;
    ge[n + p] = 0.0;
    ge[2 * n + p] = 0.0;
    ge[3 * n + p] = v;
    ge[4 * n + p] = 0.0;
    ge[5 * n + p] = v;
}

__device__ __forceinline__ double lerp_s(double a, double b, double s)
{
    return a + (b - a) * ((1.0 + s) * 0.5);
}

// Deformed box: trilinear map of each element's 8 (perturbed) vertices,
// g = w_i w_j w_k detJ J^-1 J^-T.  Not in the reference; reduces exactly
// to the brick factors when the vertices are the undisturbed width-2 grid
// (all edge differences equal -> J = I exactly).
__global__ void k_deformed_g(double *__restrict__ g, const double *__restrict__ V,
                             const double *__restrict__ zgll, const double *__restrict__ wgll,
                             const int32_t *__restrict__ lglel, int nx, int nelx, int nely,
                             int64_t nelt)
{
    const int64_t nxyz = int64_t(nx) * nx * nx;
    const int64_t gi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gi >= nxyz * nelt) return;
    const int64_t e = gi / nxyz;
    const int p = int(gi - e * nxyz);
    const int i = p % nx, j = (p / nx) % nx, k = p / (nx * nx);
    const int eg = lglel[e], nelxy = nelx * nely;
    const int ez = (eg - 1) / nelxy;
    const int ey = ((eg - 1) % nelxy) / nelx;
    const int ex = (eg - 1) % nelx;
    const int vx = nelx + 1, vy = nely + 1;
    auto X = [&](int cc, int bb, int aa, int d) {
        const int64_t v = (ex + aa) + int64_t(vx) * ((ey + bb) + int64_t(vy) * (ez + cc));
        return V[3 * v + d];
    };
    const double r = zgll[i], s = zgll[j], t = zgll[k];
    double J[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double e00 = X(0, 0, 1, d) - X(0, 0, 0, d), e10 = X(0, 1, 1, d) - X(0, 1, 0, d);
        const double e01 = X(1, 0, 1, d) - X(1, 0, 0, d), e11 = X(1, 1, 1, d) - X(1, 1, 0, d);
        J[d][0] = 0.5 * lerp_s(lerp_s(e00, e10, s), lerp_s(e01, e11, s), t);
        const double f00 = X(0, 1, 0, d) - X(0, 0, 0, d), f10 = X(0, 1, 1, d) - X(0, 0, 1, d);
        const double f01 = X(1, 1, 0, d) - X(1, 0, 0, d), f11 = X(1, 1, 1, d) - X(1, 0, 1, d);
        J[d][1] = 0.5 * lerp_s(lerp_s(f00, f10, r), lerp_s(f01, f11, r), t);
        const double h00 = X(1, 0, 0, d) - X(0, 0, 0, d), h10 = X(1, 0, 1, d) - X(0, 0, 1, d);
        const double h01 = X(1, 1, 0, d) - X(0, 1, 0, d), h11 = X(1, 1, 1, d) - X(0, 1, 1, d);
        J[d][2] = 0.5 * lerp_s(lerp_s(h00, h10, r), lerp_s(h01, h11, r), s);
    }
    const double xr = J[0][0], xs = J[0][1], xt = J[0][2];
    const double yr = J[1][0], ys = J[1][1], yt = J[1][2];
    const double zr = J[2][0], zs = J[2][1], zt = J[2][2];
    const double C00 = ys * zt - yt * zs, C10 = xt * zs - xs * zt, C20 = xs * yt - xt * ys;
    const double C01 = yt * zr - yr * zt, C11 = xr * zt - xt * zr, C21 = xt * yr - xr * yt;
    const double C02 = yr * zs - ys * zr, C12 = xs * zr - xr * zs, C22 = xr * ys - xs * yr;
    const double det = (xr * C00 + xs * C01) + xt * C02;
    const double w3 = wgll[i] * wgll[j] * wgll[k];
    const int64_t n = nxyz * nelt;
    double *ge = g + e * nxyz;
    ge[p] = (w3 * ((C00 * C00 + C10 * C10) + C20 * C20)) / det;
    ge[n + p] = (w3 * ((C00 * C01 + C10 * C11) + C20 * C21)) / det;
    ge[2 * n + p] = (w3 * ((C00 * C02 + C10 * C12) + C20 * C22)) / det;
    ge[3 * n + p] = (w3 * ((C01 * C01 + C11 * C11) + C21 * C21)) / det;
    ge[4 * n + p] = (w3 * ((C01 * C02 + C11 * C12) + C21 * C22)) / det;
    ge[5 * n + p] = (w3 * ((C02 * C02 + C12 * C12) + C22 * C22)) / det;
}

// Local diagonal of A_L (Jacobi; not in the reference):
//   t1 = sum_m (D(m,i)^2) g1(m,j,k), t2 = sum_m (D(m,j)^2) g4(i,m,k),
//   t3 = sum_m (D(m,k)^2) g6(i,j,m), m ascending;
//   d = ((((t1+t2)+t3) + 2 D(i,i)D(j,j) g2) + 2 D(i,i)D(k,k) g3)
//       + 2 D(j,j)D(k,k) g5
// (output in device point order)
__global__ void k_jacobi_local(double *__restrict__ dl, const double *__restrict__ g,
                               const double *__restrict__ D, const int32_t *__restrict__ perm,
                               int nx, int64_t nelt)
{
    const int64_t nxyz = int64_t(nx) * nx * nx;
    const int64_t gi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gi >= nxyz * nelt) return;
    const int64_t e = gi / nxyz;
    const int p = int(gi - e * nxyz);
    const int i = p % nx, j = (p / nx) % nx, k = p / (nx * nx);
    const int64_t n = nxyz * nelt;  // factor plane stride
    const double *ge = g + e * nxyz;
    const int nx2 = nx * nx;
    double t1 = (D[0 + nx * i] * D[0 + nx * i]) * ge[0 + nx * j + nx2 * k];
    double t2 = (D[0 + nx * j] * D[0 + nx * j]) * ge[3 * n + i + nx2 * k];
    double t3 = (D[0 + nx * k] * D[0 + nx * k]) * ge[5 * n + i + nx * j];
    for (int m = 1; m < nx; ++m) {
        t1 = t1 + (D[m + nx * i] * D[m + nx * i]) * ge[m + nx * j + nx2 * k];
        t2 = t2 + (D[m + nx * j] * D[m + nx * j]) * ge[3 * n + i + nx * m + nx2 * k];
        t3 = t3 + (D[m + nx * k] * D[m + nx * k]) * ge[5 * n + i + nx * j + nx2 * m];
    }
    const double dii = D[i + nx * i], djj = D[j + nx * j], dkk = D[k + nx * k];
    const double x2 = (2.0 * (dii * djj)) * ge[n + p];
    const double x3 = (2.0 * (dii * dkk)) * ge[2 * n + p];
    const double x5 = (2.0 * (djj * dkk)) * ge[4 * n + p];
    dl[e * nxyz + perm[p]] = ((((t1 + t2) + t3) + x2) + x3) + x5;
}

// Reference order <-> device order of a per-point vector.
__global__ void k_perm(double *__restrict__ dst, const double *__restrict__ src,
                       const int32_t *__restrict__ perm, int64_t nxyz, int64_t n, int to_dev)
{
    const int64_t gi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gi >= n) return;
    const int64_t e = gi / nxyz;
    const int64_t d = e * nxyz + perm[gi - e * nxyz];
    if (to_dev)
        dst[d] = src[gi];
    else
        dst[gi] = src[d];
}

__global__ void k_jacobi_finish(double *__restrict__ d, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) d[i] = 1.0 / (d[i] + 0.1);
}

// Counter-based splitmix64, value k = mix(seed + (k+1)*golden) -> U[-1,1).
__global__ void k_rand(double *__restrict__ dst, int64_t count, uint64_t seed)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint64_t z = seed + uint64_t(i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    const double u = double(z >> 11) * 0x1.0p-53;
    dst[i] = 2.0 * u - 1.0;
}

__global__ void k_col2(double *__restrict__ a, const double *__restrict__ b, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = a[i] * b[i];
}

// AoS g(6,n) (reference layout) <-> factor planes g[c][e][p].
__global__ void k_aos_soa(const double *__restrict__ src, double *__restrict__ dst,
                          int64_t nxyz, int64_t n, int dir)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 6 * n) return;
    const int64_t l = i / 6;
    const int c = int(i - l * 6);
    const int64_t s = c * n + l;  // factor planes: g[c][e][p]
    if (dir == 0)
        dst[s] = src[i];
    else
        dst[i] = src[s];
}

int grid1d(int64_t n, int threads) { return int((n + threads - 1) / threads); }

}  // namespace

GsArgs gs_args_map(const GsMap &M)
{
    GsArgs G{M.d_m2, M.d_m4, M.d_m8, M.d_roff, M.d_ridx, M.n2, M.n4, M.n8, M.nr, 0, 0, 0, 0, 0,
             M.d_m2r, M.d_m4r, M.r2, M.r4, M.n2r, M.n4r};
    auto nb = [](int64_t n) { return int((n + kRedThreads - 1) / kRedThreads); };
    G.b2 = int((M.n2 + kPairs * kRedThreads - 1) / (kPairs * kRedThreads));  // kPairs per thread
    G.b4 = nb(M.n4);
    G.b8 = nb(M.n8);
    G.br = nb(M.nr);
    return G;
}

GsArgs gs_args(const nekbone_problem *P) { return gs_args_map(P->gs); }

int reduce_blocks(int64_t n)
{
    int64_t b = (n + int64_t(kRedThreads) * 8 - 1) / (int64_t(kRedThreads) * 8);
    if (b < 1) b = 1;
    if (b > kMaxRedBlocks) b = kMaxRedBlocks;
    return int(b);
}

int gs_blocks(const nekbone_problem *P)
{
    const GsArgs G = gs_args(P);
    return G.b2 + G.b4 + G.b8 + G.br;
}

int launch_gs(nekbone_problem *P, double *u, const CgScalars *guard)
{
    const int nb = gs_blocks(P);
    if (nb == 0) return NEKBONE_OK;
    if (int rc = timed_begin(P, 1)) return rc;
    hipLaunchKernelGGL((k_gs_cls<double, false>), dim3(nb), dim3(kRedThreads), 0, stream(), u,
                       static_cast<const double *>(nullptr), static_cast<const uint8_t *>(nullptr),
                       0, gs_args(P), static_cast<double *>(nullptr), guard);
    NB_HIP(hipGetLastError());
    return timed_end(P, 1);
}

int launch_gs_cg(nekbone_problem *P, double *w, const double *p, int mbit, double *part,
                 int rev, const double *pre, int64_t npre)
{
    const int nb = gs_blocks(P);
    if (nb == 0) return NEKBONE_OK;
    if (pre && int64_t(nb) * kRedThreads < npre)
        return fail(NEKBONE_ERR_STATE, "gs: too few workgroups to fold the partials");
    if (int rc = timed_begin(P, 1)) return rc;
    GsArgs G = gs_args(P);
    G.rev = rev;
    G.pre = pre;
    G.npre = npre;
    hipLaunchKernelGGL((k_gs_cls<double, true>), dim3(nb), dim3(kRedThreads), 0, stream(), w, p,
                       P->d_flag, mbit, G, part, static_cast<const CgScalars *>(P->d_scal));
    NB_HIP(hipGetLastError());
    return timed_end(P, 1);
}

GsCgArgs gs_cg_args(nekbone_problem *P, double *w, const double *p, int mbit, double *part)
{
    return GsCgArgs{w, p, P->d_flag, mbit, gs_args(P), part, P->d_scal};
}

int launch_gs_cg_batch(const GsCgArgs *d_args, const int *d_start, int nprob, int nblocks)
{
    if (nblocks == 0) return NEKBONE_OK;
    hipLaunchKernelGGL(k_gs_cg_batch, dim3(nblocks), dim3(kRedThreads), 0, stream(), d_args,
                       d_start, nprob);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int launch_gs_map(const GsMap &M, double *u, const CgScalars *guard)
{
    const GsArgs G = gs_args_map(M);
    const int nb = G.b2 + G.b4 + G.b8 + G.br;
    if (nb == 0) return NEKBONE_OK;
    hipLaunchKernelGGL((k_gs_cls<double, false>), dim3(nb), dim3(kRedThreads), 0, stream(), u,
                       static_cast<const double *>(nullptr), static_cast<const uint8_t *>(nullptr),
                       0, G, static_cast<double *>(nullptr), guard);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

// Device form of a group list (offsets, members in the reference's
// order): members stored as memory positions e*n3 + pos[p] (pos null:
// identity), split by size class.
int gs_map_upload(GsMap &G, const std::vector<int32_t> &off, const std::vector<int32_t> &idx,
                  const int32_t *pos, int n3)
{
    G.ngroups = int64_t(off.size()) - 1;
    G.nshared = int64_t(idx.size());
    std::vector<int32_t> m2, m4, m8, roff{0}, ridx;
    auto at = [&](int32_t l) -> int32_t {
        if (!pos) return l;
        const int32_t e = l / n3;
        return e * n3 + pos[l - e * n3];
    };
    for (size_t g = 0; g + 1 < off.size(); ++g) {
        const int sz = off[g + 1] - off[g];
        std::vector<int32_t> &dst = sz == 2 ? m2 : sz == 4 ? m4 : sz == 8 ? m8 : ridx;
        for (int32_t q = off[g]; q < off[g + 1]; ++q) dst.push_back(at(idx[size_t(q)]));
        if (sz != 2 && sz != 4 && sz != 8) roff.push_back(int32_t(ridx.size()));
    }
    G.n2 = int64_t(m2.size()) / 2;
    G.n4 = int64_t(m4.size()) / 4;
    G.n8 = int64_t(m8.size()) / 8;
    G.nr = int64_t(roff.size()) - 1;
    auto up = [](int32_t **d, const std::vector<int32_t> &h) -> int {
        if (int rc = dalloc(d, int64_t(h.size()))) return rc;
        return h2d(*d, h.data(), sizeof(int32_t) * h.size());
    };
    for (auto pr : {std::make_pair(&G.d_m2, &m2), std::make_pair(&G.d_m4, &m4),
                    std::make_pair(&G.d_m8, &m8), std::make_pair(&G.d_roff, &roff),
                    std::make_pair(&G.d_ridx, &ridx)})
        if (int rc = up(pr.first, *pr.second)) return rc;
    return NEKBONE_OK;
}

void gs_map_free(GsMap &G)
{
    dfree(G.d_m2r);
    dfree(G.d_m4r);
    G.r2 = G.r4 = 0;
    G.n2r = G.n4r = 0;
    dfree(G.d_m2);
    dfree(G.d_m4);
    dfree(G.d_m8);
    dfree(G.d_roff);
    dfree(G.d_ridx);
}

int launch_gs_cg32(nekbone_problem *P, float *w, const float *p, int mbit, double *part,
                   const double *pre, int64_t npre)
{
    const int nb = gs_blocks(P);
    if (nb == 0) return NEKBONE_OK;
    if (pre && int64_t(nb) * kRedThreads < npre)
        return fail(NEKBONE_ERR_STATE, "gs: too few workgroups to fold the partials");
    if (int rc = timed_begin(P, 1)) return rc;
    GsArgs G = gs_args(P);
    G.pre = pre;
    G.npre = npre;
    hipLaunchKernelGGL((k_gs_cls<float, true>), dim3(nb), dim3(kRedThreads), 0, stream(), w, p,
                       P->d_flag, mbit, G, part, static_cast<const CgScalars *>(P->d_scal));
    NB_HIP(hipGetLastError());
    return timed_end(P, 1);
}

int launch_gs32(nekbone_problem *P, float *u)
{
    const int nb = gs_blocks(P);
    if (nb == 0) return NEKBONE_OK;
    hipLaunchKernelGGL((k_gs_cls<float, false>), dim3(nb), dim3(kRedThreads), 0, stream(), u,
                       static_cast<const float *>(nullptr), static_cast<const uint8_t *>(nullptr),
                       0, gs_args(P), static_cast<double *>(nullptr),
                       static_cast<const CgScalars *>(nullptr));
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int launch_post(nekbone_problem *P, double *w, const double *p, int mbit, bool pap)
{
    const int nb = reduce_blocks(P->n);
    if (int rc = timed_begin(P, 2)) return rc;
    if (pap)
        hipLaunchKernelGGL(k_post<true>, dim3(nb), dim3(kRedThreads), 0, stream(), w, p, P->d_c,
                           P->d_flag, mbit, P->n, P->d_part, P->d_scal);
    else
        hipLaunchKernelGGL(k_post<false>, dim3(nb), dim3(kRedThreads), 0, stream(), w, p,
                           P->d_c, P->d_flag, mbit, P->n, P->d_part,
                           static_cast<const CgScalars *>(nullptr));
    NB_HIP(hipGetLastError());
    return timed_end(P, 2);
}

int launch_mask(nekbone_problem *P, double *w, int mbit)
{
    hipLaunchKernelGGL(k_mask_flag, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(), w,
                       P->d_flag, mbit, P->n);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int launch_glsc3_raw(const double *a, const double *b, const double *m, int64_t n,
                     double *part, double *d_out)
{
    const int nb = reduce_blocks(n);
    hipLaunchKernelGGL(k_glsc3_part, dim3(nb), dim3(kRedThreads), 0, stream(), a, b, m, n,
                       part);
    hipLaunchKernelGGL(k_final_scalar, dim3(1), dim3(kRedThreads), 0, stream(),
                       static_cast<const double *>(part), nb, d_out);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int launch_glsc3(nekbone_problem *P, const double *a, const double *b, const double *m,
                 double *d_out)
{
    return launch_glsc3_raw(a, b, m, P->n, P->d_part, d_out);
}

int launch_brick_g(nekbone_problem *P)
{
    double *d_w = nullptr;
    NB_HIP(hipMallocAsync(&d_w, sizeof(double) * P->nx1, stream()));
    NB_HIP(hipMemcpyAsync(d_w, P->wgll, sizeof(double) * P->nx1, hipMemcpyHostToDevice,
                          stream()));
    hipLaunchKernelGGL(k_brick_g, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(), P->d_g, d_w,
                       P->nx1, int64_t(P->nelt));
    NB_HIP(hipGetLastError());
    NB_HIP(hipFreeAsync(d_w, stream()));
    return NEKBONE_OK;
}

int launch_deformed_g(nekbone_problem *P, const double *d_vert)
{
    double *d_zw = nullptr;
    int32_t *d_lg = nullptr;
    NB_HIP(hipMallocAsync(&d_zw, sizeof(double) * 2 * P->nx1, stream()));
    NB_HIP(hipMallocAsync(&d_lg, sizeof(int32_t) * P->nelt, stream()));
    NB_HIP(hipMemcpyAsync(d_zw, P->zgll, sizeof(double) * P->nx1, hipMemcpyHostToDevice,
                          stream()));
    NB_HIP(hipMemcpyAsync(d_zw + P->nx1, P->wgll, sizeof(double) * P->nx1,
                          hipMemcpyHostToDevice, stream()));
    NB_HIP(hipMemcpyAsync(d_lg, P->lglel.data(), sizeof(int32_t) * P->nelt,
                          hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL(k_deformed_g, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(), P->d_g,
                       d_vert, d_zw, d_zw + P->nx1, d_lg, P->nx1, P->nelx, P->nely,
                       int64_t(P->nelt));
    NB_HIP(hipGetLastError());
    NB_HIP(hipFreeAsync(d_zw, stream()));
    NB_HIP(hipFreeAsync(d_lg, stream()));
    return NEKBONE_OK;
}

int launch_jacobi_local(nekbone_problem *P, double *dl)
{
    hipLaunchKernelGGL(k_jacobi_local, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(), dl,
                       P->d_g, P->d_D, P->d_perm, P->nx1, int64_t(P->nelt));
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int launch_to_dev(nekbone_problem *P, double *dst, const double *src)
{
    const int64_t nxyz = int64_t(P->nx1) * P->nx1 * P->nx1;
    hipLaunchKernelGGL(k_perm, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(), dst, src,
                       P->d_perm, nxyz, P->n, 1);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int launch_to_ref(nekbone_problem *P, double *dst, const double *src)
{
    const int64_t nxyz = int64_t(P->nx1) * P->nx1 * P->nx1;
    hipLaunchKernelGGL(k_perm, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(), dst, src,
                       P->d_perm, nxyz, P->n, 0);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int launch_jacobi_finish(nekbone_problem *P, double *d)
{
    hipLaunchKernelGGL(k_jacobi_finish, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(), d,
                       P->n);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int launch_rand(double *dst, int64_t count, uint64_t seed)
{
    if (count <= 0) return NEKBONE_OK;
    hipLaunchKernelGGL(k_rand, dim3(grid1d(count, 256)), dim3(256), 0, stream(), dst, count,
                       seed);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int launch_col2(nekbone_problem *P, double *a, const double *b)
{
    hipLaunchKernelGGL(k_col2, dim3(grid1d(P->n, 256)), dim3(256), 0, stream(), a, b, P->n);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int launch_aos_to_soa(nekbone_problem *P, const double *aos, double *soa, int dir)
{
    const int64_t nxyz = int64_t(P->nx1) * P->nx1 * P->nx1;
    hipLaunchKernelGGL(k_aos_soa, dim3(grid1d(6 * P->n, 256)), dim3(256), 0, stream(),
                       dir == 0 ? aos : soa, dir == 0 ? soa : const_cast<double *>(aos), nxyz,
                       P->n, dir);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

}  // namespace nb
