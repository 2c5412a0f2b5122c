// Gather-scatter, fused CG vector kernels, deterministic reductions and
// set-up kernels.  All FP64; the library is compiled with
// -ffp-contract=off so each expression rounds exactly as written.
#include "nb_internal.h"

namespace nb {
namespace {

constexpr int kRedThreads = 256;
constexpr int kMaxRedBlocks = 1024;

// Fixed-shape block tree reduction of two values (deterministic: the
// combination order depends only on thread indices).
__device__ __forceinline__ void block_sum2(double &v0, double &v1)
{
    __shared__ double s0[kRedThreads], s1[kRedThreads];
    const int t = threadIdx.x;
    s0[t] = v0;
    s1[t] = v1;
    __syncthreads();
#pragma unroll
    for (int h = kRedThreads / 2; h > 0; h >>= 1) {
        if (t < h) {
            s0[t] = s0[t] + s0[t + h];
            s1[t] = s1[t] + s1[t + h];
        }
        __syncthreads();
    }
    v0 = s0[0];
    v1 = s1[0];
}

// ---------------------------------------------------------------------
// dssum: per shared group, t = u[i0] + u[i1] + ... in ascending local
// index, then t written to every member (src/jl/gs_local.c:60-87).
// ---------------------------------------------------------------------
__global__ void k_gs(double *__restrict__ u, const int32_t *__restrict__ off,
                     const int32_t *__restrict__ idx, int64_t ngroups,
                     const CgScalars *__restrict__ scal)
{
    if (scal && scal->done) return;
    const int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= ngroups) return;
    const int a = off[gid], b = off[gid + 1];
    double t = u[idx[a]];
    for (int q = a + 1; q < b; ++q) t += u[idx[q]];
    for (int q = a; q < b; ++q) u[idx[q]] = t;
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

// Sum of nblk partials (two interleaved streams) in a fixed order.
__device__ __forceinline__ void sum_partials(const double *__restrict__ part, int nblk,
                                             double &s0, double &s1)
{
    double a0 = 0.0, a1 = 0.0;
    for (int i = threadIdx.x; i < nblk; i += kRedThreads) {
        a0 = a0 + part[i];
        a1 = a1 + part[kMaxRedBlocks + i];
    }
    block_sum2(a0, a1);
    s0 = a0;
    s1 = a1;
}

__global__ void __launch_bounds__(kRedThreads)
    k_final_scalar(const double *__restrict__ part, int nblk, double *__restrict__ out)
{
    double s0, s1;
    sum_partials(part, nblk, s0, s1);
    if (threadIdx.x == 0) *out = s0;
}

// ---------------------------------------------------------------------
// CG scalar bookkeeping (src/cg.f:44-66), all on the device.
// ---------------------------------------------------------------------
__device__ void prepare_next(CgScalars *s, double *hist, int hcap)
{
    s->rtz2 = s->rtz1;
    s->rtz1 = s->rtz_next;
    s->beta = (s->iter == 0) ? 0.0 : s->rtz1 / s->rtz2;
    const int h = 1 + 3 * s->iter;
    if (h < hcap) hist[h] = s->rtz1;
}

// x = 0, r = mask(f); partials of glsc3(r,c,r) and glsc3(r,c,z(r)).
template <int PRE>
__global__ void __launch_bounds__(kRedThreads)
    k_cg_init(double *__restrict__ x, double *__restrict__ r, const double *__restrict__ f,
              const double *__restrict__ c, const double *__restrict__ dinv,
              const uint8_t *__restrict__ flag, int mbit, int64_t n, double *__restrict__ part)
{
    double a0 = 0.0, a1 = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kRedThreads;
    for (int64_t i = int64_t(blockIdx.x) * kRedThreads + threadIdx.x; i < n; i += stride) {
        x[i] = 0.0;
        const double v = (flag[i] & mbit) ? 0.0 : f[i];
        r[i] = v;
        const double rc = v * c[i];
        a0 = a0 + rc * v;
        if constexpr (PRE == NEKBONE_PRECOND_JACOBI) a1 = a1 + rc * (v * dinv[i]);
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a0;
        part[kMaxRedBlocks + blockIdx.x] = a1;
    }
}

template <int PRE>
__global__ void __launch_bounds__(kRedThreads)
    k_cg_init_final(const double *__restrict__ part, int nblk, CgScalars *__restrict__ s,
                    double *__restrict__ hist, int hcap, double tol2)
{
    double rtr, rtz;
    sum_partials(part, nblk, rtr, rtz);
    if (threadIdx.x == 0) {
        s->rtr0 = rtr;
        s->rtr = rtr;
        s->rtz1 = 1.0;
        s->rtz_next = (PRE == NEKBONE_PRECOND_JACOBI) ? rtz : rtr;
        s->iter = 0;
        s->done = (tol2 > 0.0 && rtr == 0.0) ? 1 : 0;
        s->status = 0;
        s->tol2 = tol2;
        if (hcap > 0) hist[0] = rtr;
        prepare_next(s, hist, hcap);
    }
}

__global__ void __launch_bounds__(kRedThreads)
    k_pap_final(const double *__restrict__ part, int nblk, CgScalars *__restrict__ s,
                double *__restrict__ hist, int hcap)
{
    if (s->done) return;
    double pap, dummy;
    sum_partials(part, nblk, pap, dummy);
    if (threadIdx.x == 0) {
        s->pap = pap;
        s->alpha = s->rtz1 / pap;
        if (!(pap > 0.0) || !isfinite(pap)) {
            s->status = 1;
            if (s->tol2 > 0.0) s->done = 1;
        }
        const int h = 2 + 3 * s->iter;
        if (h < hcap) hist[h] = pap;
    }
}

// x += alpha p; r += (-alpha) w (add2s2 x2, src/cg.f:61-64); partials of
// rtr = glsc3(r,c,r) (src/cg.f:66) and of the next glsc3(r,c,z).
template <int PRE, bool FMA>
__global__ void __launch_bounds__(kRedThreads)
    k_cg_update(double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
                const double *__restrict__ w, const double *__restrict__ c,
                const double *__restrict__ dinv, int64_t n, double *__restrict__ part,
                const CgScalars *__restrict__ s)
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
        part[kMaxRedBlocks + blockIdx.x] = a1;
    }
}

template <int PRE>
__global__ void __launch_bounds__(kRedThreads)
    k_update_final(const double *__restrict__ part, int nblk, CgScalars *__restrict__ s,
                   double *__restrict__ hist, int hcap)
{
    if (s->done) return;
    double rtr, rtz;
    sum_partials(part, nblk, rtr, rtz);
    if (threadIdx.x == 0) {
        s->rtr = rtr;
        const int h = 3 + 3 * s->iter;
        if (h < hcap) hist[h] = rtr;
        s->iter += 1;
        s->rtz_next = (PRE == NEKBONE_PRECOND_JACOBI) ? rtz : rtr;
        if (s->tol2 > 0.0 && rtr <= s->tol2 * s->rtr0) s->done = 1;
        prepare_next(s, hist, hcap);
    }
}

// ---------------------------------------------------------------------
// Set-up kernels
// ---------------------------------------------------------------------

// Brick factors g1=g4=g6=w_i w_j w_k, others 0 (src/prox_setup.f:27-52),
// element-blocked SoA.
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
    double *ge = g + e * 6 * nxyz;
    ge[p] = v;
    ge[nxyz + p] = 0.0;
    ge[2 * nxyz + p] = 0.0;
    ge[3 * nxyz + p] = v;
    ge[4 * nxyz + p] = 0.0;
    ge[5 * nxyz + p] = v;
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
    double *ge = g + e * 6 * nxyz;
    ge[p] = (w3 * ((C00 * C00 + C10 * C10) + C20 * C20)) / det;
    ge[nxyz + p] = (w3 * ((C00 * C01 + C10 * C11) + C20 * C21)) / det;
    ge[2 * nxyz + p] = (w3 * ((C00 * C02 + C10 * C12) + C20 * C22)) / det;
    ge[3 * nxyz + p] = (w3 * ((C01 * C01 + C11 * C11) + C21 * C21)) / det;
    ge[4 * nxyz + p] = (w3 * ((C01 * C02 + C11 * C12) + C21 * C22)) / det;
    ge[5 * nxyz + p] = (w3 * ((C02 * C02 + C12 * C12) + C22 * C22)) / det;
}

// Local diagonal of A_L (Jacobi; not in the reference):
//   t1 = sum_m (D(m,i)^2) g1(m,j,k), t2 = sum_m (D(m,j)^2) g4(i,m,k),
//   t3 = sum_m (D(m,k)^2) g6(i,j,m), m ascending;
//   d = ((((t1+t2)+t3) + 2 D(i,i)D(j,j) g2) + 2 D(i,i)D(k,k) g3)
//       + 2 D(j,j)D(k,k) g5
__global__ void k_jacobi_local(double *__restrict__ dl, const double *__restrict__ g,
                               const double *__restrict__ D, int nx, int64_t nelt)
{
    const int64_t nxyz = int64_t(nx) * nx * nx;
    const int64_t gi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gi >= nxyz * nelt) return;
    const int64_t e = gi / nxyz;
    const int p = int(gi - e * nxyz);
    const int i = p % nx, j = (p / nx) % nx, k = p / (nx * nx);
    const double *ge = g + e * 6 * nxyz;
    const int nx2 = nx * nx;
    double t1 = (D[0 + nx * i] * D[0 + nx * i]) * ge[0 + nx * j + nx2 * k];
    double t2 = (D[0 + nx * j] * D[0 + nx * j]) * ge[3 * nxyz + i + nx2 * k];
    double t3 = (D[0 + nx * k] * D[0 + nx * k]) * ge[5 * nxyz + i + nx * j];
    for (int m = 1; m < nx; ++m) {
        t1 = t1 + (D[m + nx * i] * D[m + nx * i]) * ge[m + nx * j + nx2 * k];
        t2 = t2 + (D[m + nx * j] * D[m + nx * j]) * ge[3 * nxyz + i + nx * m + nx2 * k];
        t3 = t3 + (D[m + nx * k] * D[m + nx * k]) * ge[5 * nxyz + i + nx * j + nx2 * m];
    }
    const double dii = D[i + nx * i], djj = D[j + nx * j], dkk = D[k + nx * k];
    const double x2 = (2.0 * (dii * djj)) * ge[nxyz + p];
    const double x3 = (2.0 * (dii * dkk)) * ge[2 * nxyz + p];
    const double x5 = (2.0 * (djj * dkk)) * ge[4 * nxyz + p];
    dl[gi] = ((((t1 + t2) + t3) + x2) + x3) + x5;
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

// AoS g(6,n) (reference layout) <-> element-blocked SoA g[e][c][p].
__global__ void k_aos_soa(const double *__restrict__ src, double *__restrict__ dst,
                          int64_t nxyz, int64_t n, int dir)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 6 * n) return;
    const int64_t l = i / 6;
    const int c = int(i - l * 6);
    const int64_t e = l / nxyz, p = l - e * nxyz;
    const int64_t s = (e * 6 + c) * nxyz + p;
    if (dir == 0)
        dst[s] = src[i];
    else
        dst[i] = src[s];
}

int grid1d(int64_t n, int threads) { return int((n + threads - 1) / threads); }

}  // namespace

int reduce_blocks(int64_t n)
{
    int64_t b = (n + int64_t(kRedThreads) * 8 - 1) / (int64_t(kRedThreads) * 8);
    if (b < 1) b = 1;
    if (b > kMaxRedBlocks) b = kMaxRedBlocks;
    return int(b);
}

int launch_gs(nekbone_problem *P, double *u, const CgScalars *guard)
{
    if (P->gs.ngroups == 0) return NEKBONE_OK;
    if (int rc = timed_begin(P, 1)) return rc;
    hipLaunchKernelGGL(k_gs, dim3(grid1d(P->gs.ngroups, 256)), dim3(256), 0, stream(), u,
                       P->gs.d_off, P->gs.d_idx, P->gs.ngroups, guard);
    NB_HIP(hipGetLastError());
    return timed_end(P, 1);
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

int launch_glsc3(nekbone_problem *P, const double *a, const double *b, const double *m,
                 double *d_out)
{
    const int nb = reduce_blocks(P->n);
    hipLaunchKernelGGL(k_glsc3_part, dim3(nb), dim3(kRedThreads), 0, stream(), a, b, m, P->n,
                       P->d_part);
    hipLaunchKernelGGL(k_final_scalar, dim3(1), dim3(kRedThreads), 0, stream(), P->d_part, nb,
                       d_out);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
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
                       P->d_g, P->d_D, P->nx1, int64_t(P->nelt));
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

int launch_cg_init(nekbone_problem *P, double *x, double *r, const double *f, int mbit,
                   int precond, double tol2)
{
    const int nb = reduce_blocks(P->n);
    if (precond == NEKBONE_PRECOND_JACOBI) {
        hipLaunchKernelGGL(k_cg_init<NEKBONE_PRECOND_JACOBI>, dim3(nb), dim3(kRedThreads), 0,
                           stream(), x, r, f, P->d_c, P->d_dinv, P->d_flag, mbit, P->n,
                           P->d_part);
        hipLaunchKernelGGL(k_cg_init_final<NEKBONE_PRECOND_JACOBI>, dim3(1), dim3(kRedThreads),
                           0, stream(), P->d_part, nb, P->d_scal, P->d_hist, P->hist_cap, tol2);
    } else {
        hipLaunchKernelGGL(k_cg_init<NEKBONE_PRECOND_IDENTITY>, dim3(nb), dim3(kRedThreads), 0,
                           stream(), x, r, f, P->d_c, P->d_dinv, P->d_flag, mbit, P->n,
                           P->d_part);
        hipLaunchKernelGGL(k_cg_init_final<NEKBONE_PRECOND_IDENTITY>, dim3(1),
                           dim3(kRedThreads), 0, stream(), P->d_part, nb, P->d_scal, P->d_hist,
                           P->hist_cap, tol2);
    }
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

// One PCG iteration (src/cg.f:48-75) as 6 launches, no host round trip:
// ax (with the fused p update) -> dssum -> shift+mask+pap -> alpha ->
// x/r update with rtr (and rtz) -> beta.
int launch_cg_iteration(nekbone_problem *P, double *x, double *r, int precond, int mbit)
{
    const int nb = reduce_blocks(P->n);
    const bool jac = precond == NEKBONE_PRECOND_JACOBI;
    const bool fma = P->arith == NEKBONE_ARITH_FMA;
    if (int rc = launch_ax_local(P, P->d_w, nullptr, jac ? 2 : 1, P->d_p, r)) return rc;
    if (int rc = launch_gs(P, P->d_w, P->d_scal)) return rc;
    if (int rc = launch_post(P, P->d_w, P->d_p, mbit, true)) return rc;
    if (int rc = timed_begin(P, 4)) return rc;
    hipLaunchKernelGGL(k_pap_final, dim3(1), dim3(kRedThreads), 0, stream(), P->d_part, nb,
                       P->d_scal, P->d_hist, P->hist_cap);
    NB_HIP(hipGetLastError());
    if (int rc = timed_end(P, 4)) return rc;
    if (int rc = timed_begin(P, 3)) return rc;
#define NB_UPD(PRE, F)                                                                      \
    hipLaunchKernelGGL((k_cg_update<PRE, F>), dim3(nb), dim3(kRedThreads), 0, stream(), x, r, \
                       P->d_p, P->d_w, P->d_c, P->d_dinv, P->n, P->d_part, P->d_scal)
    if (jac) {
        if (fma)
            NB_UPD(NEKBONE_PRECOND_JACOBI, true);
        else
            NB_UPD(NEKBONE_PRECOND_JACOBI, false);
    } else {
        if (fma)
            NB_UPD(NEKBONE_PRECOND_IDENTITY, true);
        else
            NB_UPD(NEKBONE_PRECOND_IDENTITY, false);
    }
#undef NB_UPD
    NB_HIP(hipGetLastError());
    if (int rc = timed_end(P, 3)) return rc;
    if (int rc = timed_begin(P, 4)) return rc;
    if (jac)
        hipLaunchKernelGGL(k_update_final<NEKBONE_PRECOND_JACOBI>, dim3(1), dim3(kRedThreads), 0,
                           stream(), P->d_part, nb, P->d_scal, P->d_hist, P->hist_cap);
    else
        hipLaunchKernelGGL(k_update_final<NEKBONE_PRECOND_IDENTITY>, dim3(1), dim3(kRedThreads),
                           0, stream(), P->d_part, nb, P->d_scal, P->d_hist, P->hist_cap);
    NB_HIP(hipGetLastError());
    return timed_end(P, 4);
}

}  // namespace nb
