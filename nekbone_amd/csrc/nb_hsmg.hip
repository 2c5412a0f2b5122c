// HSMG preconditioner application on the device: h1mg_solve
// (src/hsmg.f:960-1034), an additive two-level Schwarz/FDM V-cycle with a
// trivial coarse solve (QQ^T on the order-1 level).
//
// Per Schwarz level (h1mg_schwarz, :1036-1095, sigma = 1):
//   k_fdm     mask r in place (the reference's side effect, :1067), build
//             the extended element (own points, ghost face = the
//             neighbour's point one inside, as the (n+2)^3 exchange of
//             :1071-1076 delivers it), local solve e = S D^-1 S^T r with
//             S = S_z x S_y x S_x (h1mg_do_fast, :1297-1320); writes the
//             own part and the ghost faces of e
//   k_ovl     add the neighbours' ghost results onto the first interior
//             layer (:1081-1085)
//   gs        dssum on the level (:1089)
//   k_mask_w  mask, then the square-rooted Schwarz weights (:1090, :1144-1178)
// Restriction k_rstr (h1mg_rstr, :1340-1357) and interpolation-add
// k_intp (h1mg_intp, :1436-1443) are per-element tensor products through
// LDS.  Every sum keeps the reference's order (mxm: left to right from the
// first product; the ghost exchange as (own + neighbour) - own).
#include <vector>

#include "nb_device.h"

namespace nb {
namespace {

constexpr int kT = 256;

struct FdmArgs {
    double *r;
    const uint8_t *bnd;
    const int32_t *pos, *gsrc, *gin;
    const double *ops;
    const int16_t *eop;
    const double *eps;
    double *eown, *ghost;
    int nh, nl;
};

__device__ __forceinline__ void ghost_ext(int f, int t, int nh, int &I, int &J, int &K)
{
    const int nl = nh + 2, a = t % nh + 1, b = t / nh + 1;
    const int side = (f & 1) ? nl - 1 : 0;
    if (f < 2) {
        I = side;
        J = a;
        K = b;
    } else if (f < 4) {
        I = a;
        J = side;
        K = b;
    } else {
        I = a;
        J = b;
        K = side;
    }
}

// out(i,jk) = sum_m M(i,m) in(m,jk)       (first index)
__device__ void tx(double *out, const double *in, const double *M, int nv, int nu, int n23)
{
    for (int q = threadIdx.x; q < nv * n23; q += blockDim.x) {
        const int i = q % nv, jk = q / nv;
        double s = M[i] * in[nu * jk];
        for (int m = 1; m < nu; ++m) s = s + M[i + nv * m] * in[m + nu * jk];
        out[q] = s;
    }
}
// out(i,j,k) = sum_m in(i,m,k) Bt(m,j)     (second index); in is (n1, nu, n3)
__device__ void ty(double *out, const double *in, const double *Bt, int n1, int nv, int nu, int n3)
{
    for (int q = threadIdx.x; q < n1 * nv * n3; q += blockDim.x) {
        const int i = q % n1, j = (q / n1) % nv, k = q / (n1 * nv);
        const double *row = in + n1 * nu * k + i;
        double s = row[0] * Bt[nu * j];
        for (int m = 1; m < nu; ++m) s = s + row[n1 * m] * Bt[m + nu * j];
        out[q] = s;
    }
}
// out(ij,k) = sum_m in(ij,m) Ct(m,k)       (third index); in is (n12, nu)
__device__ void tz(double *out, const double *in, const double *Ct, int n12, int nv, int nu)
{
    for (int q = threadIdx.x; q < n12 * nv; q += blockDim.x) {
        const int ij = q % n12, k = q / n12;
        double s = in[ij] * Ct[nu * k];
        for (int m = 1; m < nu; ++m) s = s + in[ij + n12 * m] * Ct[m + nu * k];
        out[q] = s;
    }
}

__global__ void __launch_bounds__(kT) k_fdm(FdmArgs a)
{
    extern __shared__ double sm[];
    const int nh = a.nh, nl = a.nl, l2 = nl * nl, l3 = l2 * nl, n3 = nh * nh * nh;
    const int h2 = nh * nh, g2 = 6 * h2, opsz = 2 * l2 + nl;
    double *A = sm, *B = sm + l3, *S = sm + 2 * l3;
    const int e = blockIdx.x, tid = threadIdx.x;
    for (int d = 0; d < 3; ++d) {
        const double *src = a.ops + int64_t(a.eop[3 * e + d]) * opsz;
        for (int q = tid; q < opsz; q += blockDim.x) S[d * opsz + q] = src[q];
    }
    for (int q = tid; q < l3; q += blockDim.x) A[q] = 0.0;
    __syncthreads();
    const int64_t base = int64_t(e) * n3;
    for (int p = tid; p < n3; p += blockDim.x) {
        const int i = p % nh, j = (p / nh) % nh, k = p / h2;
        const int64_t m = base + a.pos[p];
        double v = a.r[m];
        if (a.bnd[m]) {
            v = 0.0;
            a.r[m] = 0.0;
        }
        A[(i + 1) + nl * (j + 1) + l2 * (k + 1)] = v;
    }
    for (int q = tid; q < g2; q += blockDim.x) {
        int I, J, K;
        ghost_ext(q / h2, q % h2, nh, I, J, K);
        const int32_t s = a.gsrc[int64_t(e) * g2 + q];
        double v = 0.0;
        if (s >= 0) {
            const int32_t o = a.gin[int64_t(e) * g2 + q];
            const double vo = a.bnd[o] ? 0.0 : a.r[o];
            const double vs = a.bnd[s] ? 0.0 : a.r[s];
            v = (vo + vs) - vo;
        }
        A[I + nl * J + l2 * K] = v;
    }
    __syncthreads();
    const double *Sx = S, *SxT = S + l2, *lx = S + 2 * l2;
    const double *Sy = S + opsz, *SyT = Sy + l2, *ly = Sy + 2 * l2;
    const double *Sz = S + 2 * opsz, *SzT = Sz + l2, *lz = Sz + 2 * l2;
    tx(B, A, SxT, nl, nl, l2);
    __syncthreads();
    ty(A, B, Sy, nl, nl, nl, nl);
    __syncthreads();
    tz(B, A, Sz, l2, nl, nl);
    __syncthreads();
    const double eps = a.eps[e];
    for (int q = tid; q < l3; q += blockDim.x) {
        const int i = q % nl, j = (q / nl) % nl, k = q / l2;
        const double diag = lx[i] + ly[j] + lz[k];
        const double d = (diag > eps) ? 1.0 / diag : 0.0;
        A[q] = d * B[q];
    }
    __syncthreads();
    tx(B, A, Sx, nl, nl, l2);
    __syncthreads();
    ty(A, B, SyT, nl, nl, nl, nl);
    __syncthreads();
    tz(B, A, SzT, l2, nl, nl);
    __syncthreads();
    for (int p = tid; p < n3; p += blockDim.x) {
        const int i = p % nh, j = (p / nh) % nh, k = p / h2;
        a.eown[base + a.pos[p]] = B[(i + 1) + nl * (j + 1) + l2 * (k + 1)];
    }
    for (int q = tid; q < g2; q += blockDim.x) {
        int I, J, K;
        ghost_ext(q / h2, q % h2, nh, I, J, K);
        a.ghost[int64_t(e) * g2 + q] = B[I + nl * J + l2 * K];
    }
}

// e(first interior layer) += the neighbour's ghost result, x faces, then
// y, then z (src/hsmg.f:1081-1085); in place on eown.
__global__ void k_ovl(double *__restrict__ eown, const double *__restrict__ ghost,
                      const int32_t *__restrict__ gpart, const int32_t *__restrict__ pos, int nh,
                      int64_t n)
{
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int n3 = nh * nh * nh, h2 = nh * nh, g2 = 6 * h2;
    const int64_t e = q / n3;
    const int p = int(q - e * n3);
    const int i = p % nh, j = (p / nh) % nh, k = p / h2;
    const int64_t m = e * n3 + pos[p];
    double v = eown[m];
    const int c[3] = {i, j, k};
    const int t[3] = {j + nh * k, i + nh * k, i + nh * j};
    for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side) {
            if (c[d] != (side ? nh - 2 : 1)) continue;
            const int64_t slot = e * g2 + (2 * d + side) * h2 + t[d];
            const int32_t g = gpart[slot];
            if (g < 0) continue;
            const double own = ghost[slot];
            v = v + ((own + ghost[g]) - own);
        }
    eown[m] = v;
}

__global__ void k_mask_w(double *__restrict__ out, const double *__restrict__ in,
                         const uint8_t *__restrict__ bnd, const double *__restrict__ sw, int64_t n)
{
    const int64_t m = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double v = bnd[m] ? 0.0 : in[m];
    out[m] = v * sw[m];
}

// Restriction of one element from the finer level F to the coarser C:
// rc = [J^T x J^T x J^T] (c_F .* rf) (h1mg_do_wt then h1mg_tnsr1_3d);
// with MASK the coarse result is written masked.
template <bool MASK>
__global__ void __launch_bounds__(kT)
    k_rstr(double *__restrict__ rc, const double *__restrict__ rf, const double *__restrict__ cf,
           const int32_t *__restrict__ posf, const double *__restrict__ jh,
           const double *__restrict__ jht, const uint8_t *__restrict__ bndc, int nc, int nf)
{
    extern __shared__ double sm[];
    const int f3 = nf * nf * nf, c3 = nc * nc * nc;
    double *U = sm, *W1 = sm + f3, *W2 = W1 + nc * nf * nf;
    const int64_t e = blockIdx.x;
    for (int p = threadIdx.x; p < f3; p += blockDim.x) {
        const int64_t m = e * f3 + posf[p];
        U[p] = rf[m] * cf[m];
    }
    __syncthreads();
    tx(W1, U, jht, nc, nf, nf * nf);         // A = jht (nc x nf)
    __syncthreads();
    ty(W2, W1, jh, nc, nc, nf, nf);           // Bt = jh (nf x nc)
    __syncthreads();
    tz(U, W2, jh, nc * nc, nc, nf);           // Ct = jh
    __syncthreads();
    for (int p = threadIdx.x; p < c3; p += blockDim.x) {
        const int64_t m = e * c3 + p;
        rc[m] = (MASK && bndc[m]) ? 0.0 : U[p];
    }
}

// Interpolation-add of one element: uf += [J x J x J] uc
// (h1mg_intp then the add of h1mg_solve); with MASK the coarse input is
// masked first (the coarse solve's final mask, :1011).
template <bool MASK>
__global__ void __launch_bounds__(kT)
    k_intp(double *__restrict__ uf, const int32_t *__restrict__ posf,
           const double *__restrict__ uc, const uint8_t *__restrict__ bndc,
           const double *__restrict__ jh, const double *__restrict__ jht, int nc, int nf)
{
    extern __shared__ double sm[];
    const int f3 = nf * nf * nf, c3 = nc * nc * nc;
    double *U = sm, *W1 = sm + f3, *W2 = W1 + nf * nc * nc;
    const int64_t e = blockIdx.x;
    for (int p = threadIdx.x; p < c3; p += blockDim.x) {
        const int64_t m = e * c3 + p;
        U[p] = (MASK && bndc[m]) ? 0.0 : uc[m];
    }
    __syncthreads();
    tx(W1, U, jh, nf, nc, nc * nc);           // A = jh (nf x nc)
    __syncthreads();
    ty(W2, W1, jht, nf, nf, nc, nc);          // Bt = jht (nc x nf)
    __syncthreads();
    tz(U, W2, jht, nf * nf, nf, nc);          // Ct = jht
    __syncthreads();
    for (int p = threadIdx.x; p < f3; p += blockDim.x) {
        const int64_t m = e * f3 + posf[p];
        uf[m] = uf[m] + U[p];
    }
}

int grid(int64_t n) { return int((n + kT - 1) / kT); }

// One Schwarz level: out = W^1/2 mask(QQ^T(ovl(fdm(mask(r))))); r masked in place.
int schwarz(nekbone_problem *P, HsmgLevel &L, double *out, double *r)
{
    const int nelt = P->nelt;
    FdmArgs a{r,       L.d_bnd, L.d_pos, L.d_gsrc, L.d_gin, L.d_ops,
              L.d_eop, L.d_eps, L.d_eown, L.d_ghost, L.nh, L.nl};
    const int l3 = L.nl * L.nl * L.nl;
    const size_t lds = sizeof(double) * size_t(2 * l3 + 3 * (2 * L.nl * L.nl + L.nl));
    hipLaunchKernelGGL(k_fdm, dim3(nelt), dim3(kT), lds, stream(), a);
    NB_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_ovl, dim3(grid(L.n)), dim3(kT), 0, stream(), L.d_eown,
                       static_cast<const double *>(L.d_ghost),
                       static_cast<const int32_t *>(L.d_gpart),
                       static_cast<const int32_t *>(L.d_pos), L.nh, L.n);
    NB_HIP(hipGetLastError());
    if (L.fine) {
        if (int rc = launch_gs(P, L.d_eown, nullptr)) return rc;
    } else if (int rc = launch_gs_map(L.gs, L.d_eown)) {
        return rc;
    }
    hipLaunchKernelGGL(k_mask_w, dim3(grid(L.n)), dim3(kT), 0, stream(), out,
                       static_cast<const double *>(L.d_eown),
                       static_cast<const uint8_t *>(L.d_bnd),
                       static_cast<const double *>(L.d_sw), L.n);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int rstr(nekbone_problem *P, HsmgLevel &C, HsmgLevel &F, double *rc, const double *rf, bool mask)
{
    const int nc = C.nh, nf = F.nh;
    const size_t lds = sizeof(double) * size_t(nf * nf * nf + nc * nf * nf + nc * nc * nf);
    if (mask)
        hipLaunchKernelGGL(k_rstr<true>, dim3(P->nelt), dim3(kT), lds, stream(), rc, rf,
                           static_cast<const double *>(F.d_c),
                           static_cast<const int32_t *>(F.d_pos),
                           static_cast<const double *>(C.d_jh),
                           static_cast<const double *>(C.d_jht),
                           static_cast<const uint8_t *>(C.d_bnd), nc, nf);
    else
        hipLaunchKernelGGL(k_rstr<false>, dim3(P->nelt), dim3(kT), lds, stream(), rc, rf,
                           static_cast<const double *>(F.d_c),
                           static_cast<const int32_t *>(F.d_pos),
                           static_cast<const double *>(C.d_jh),
                           static_cast<const double *>(C.d_jht),
                           static_cast<const uint8_t *>(C.d_bnd), nc, nf);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

int intp(nekbone_problem *P, HsmgLevel &C, HsmgLevel &F, double *uf, const double *uc, bool mask)
{
    const int nc = C.nh, nf = F.nh;
    const size_t lds = sizeof(double) * size_t(nf * nf * nf + nf * nc * nc + nf * nf * nc);
    if (mask)
        hipLaunchKernelGGL(k_intp<true>, dim3(P->nelt), dim3(kT), lds, stream(), uf,
                           static_cast<const int32_t *>(F.d_pos), uc,
                           static_cast<const uint8_t *>(C.d_bnd),
                           static_cast<const double *>(C.d_jh),
                           static_cast<const double *>(C.d_jht), nc, nf);
    else
        hipLaunchKernelGGL(k_intp<false>, dim3(P->nelt), dim3(kT), lds, stream(), uf,
                           static_cast<const int32_t *>(F.d_pos), uc,
                           static_cast<const uint8_t *>(C.d_bnd),
                           static_cast<const double *>(C.d_jh),
                           static_cast<const double *>(C.d_jht), nc, nf);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

}  // namespace

// h1mg_solve (src/hsmg.f:960-1034): z = M^-1 r on the finest level
// (device point order); r is masked in place.
int launch_hsmg_apply(nekbone_problem *P, double *z, double *r)
{
    Hsmg *M = P->mg;
    if (!M) return fail(NEKBONE_ERR_STATE, "hsmg: not set up");
    const int lmax = M->lmax;
    HsmgLevel *L = M->L;
    if (int rc = timed_begin(P, 2)) return rc;
    if (int rc = schwarz(P, L[lmax], z, r)) return rc;
    const double *rcur = r;  // r := rhs, masked by the Schwarz step
    for (int l = lmax - 1; l >= 2; --l) {
        if (int rc = rstr(P, L[l], L[l + 1], L[l].d_r, rcur, false)) return rc;
        if (int rc = launch_gs_map(L[l].gs, L[l].d_r)) return rc;
        if (int rc = schwarz(P, L[l], L[l].d_e, L[l].d_r)) return rc;
        rcur = L[l].d_r;
    }
    // coarse level: e1 = mask(QQ^T mask(J^T r2)) (hsmg_coarse_solve, :945-957)
    if (int rc = rstr(P, L[1], L[2], L[1].d_e, rcur, true)) return rc;
    if (int rc = launch_gs_map(L[1].gs, L[1].d_e)) return rc;
    for (int l = 2; l <= lmax - 1; ++l)
        if (int rc = intp(P, L[l - 1], L[l], L[l].d_e, L[l - 1].d_e, l - 1 == 1)) return rc;
    if (int rc = intp(P, L[lmax - 1], L[lmax], z, L[lmax - 1].d_e, lmax - 1 == 1)) return rc;
    return timed_end(P, 2);
}

}  // namespace nb

using namespace nb;

extern "C" int nekbone_setup_hsmg(nekbone_problem *P)
{
    if (int rc = check(P)) return rc;
    if (int rc = hsmg_build(P)) {
        hsmg_free(P);
        return rc;
    }
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

extern "C" int nekbone_hsmg_solve(nekbone_problem *P, double *z, double *r)
{
    if (int rc = check(P)) return rc;
    if (!P->mg) return fail(NEKBONE_ERR_STATE, "hsmg_solve: nekbone_setup_hsmg not called");
    if (!z || !r) return fail(NEKBONE_ERR_ARG, "hsmg_solve: null array");
    if (int rc = upload_dev(P, P->d_r, r)) return rc;
    if (int rc = launch_hsmg_apply(P, P->mg->d_z, P->d_r)) return rc;
    if (int rc = download_ref(P, z, P->mg->d_z)) return rc;
    return download_ref(P, r, P->d_r);
}
