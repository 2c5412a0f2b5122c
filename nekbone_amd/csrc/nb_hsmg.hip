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
    const int16_t *eop;
    double *eown, *ghost;
    int nelt;
};

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

// Sizes of the local FDM solve on a level with NH points per direction:
// the extended element has NL = NH+2 points per direction; one thread per
// grid line (NL^2 per element), EPB elements per workgroup.
template <int NH>
struct FdmShape {
    static constexpr int NL = NH + 2, L2 = NL * NL, L3 = L2 * NL;
    static constexpr int EPB = L2 >= 256 ? 1 : 256 / L2;
    static constexpr int NT = (EPB * L2 + 63) / 64 * 64;
    static constexpr int OPS = 2 * L2 + NL;  // S, S^T (column-major), eigenvalues
};

// v <- C v for a column-major NL x NL matrix C in LDS: out_i is summed
// left to right from the first product, C(i,0) v_0 + C(i,1) v_1 + ...
// (the mxm order); FP64 rounds every product and sum, FP32 fuses them.
template <typename T, int NL>
__device__ __forceinline__ void line_apply(T (&v)[NL], const T *__restrict__ C)
{
    T o[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) o[i] = C[i] * v[0];
#pragma unroll
    for (int m = 1; m < NL; ++m)
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            if constexpr (sizeof(T) == 4)
                o[i] = __builtin_fmaf(C[i + NL * m], v[m], o[i]);
            else
                o[i] = o[i] + C[i + NL * m] * v[m];
        }
#pragma unroll
    for (int i = 0; i < NL; ++i) v[i] = o[i];
}

template <typename T, int NL>
__device__ __forceinline__ void line_load(T (&v)[NL], const T *b, int base, int stride)
{
#pragma unroll
    for (int m = 0; m < NL; ++m) v[m] = b[base + stride * m];
}

template <typename T, int NL>
__device__ __forceinline__ void line_store(const T (&v)[NL], T *b, int base, int stride)
{
#pragma unroll
    for (int m = 0; m < NL; ++m) b[base + stride * m] = v[m];
}

// Extended-block position of ghost slot t of face f (x-, x+, y-, y+, z-,
// z+): the face's (a, b) = (t % NH + 1, t / NH + 1) in its two tangential
// directions, the normal coordinate 0 or NL-1.
template <int NH>
__device__ __forceinline__ int ghost_ext(int f, int t)
{
    constexpr int NL = NH + 2;
    const int a = t % NH + 1, b = t / NH + 1, side = (f & 1) ? NL - 1 : 0;
    if (f < 2) return side + NL * a + NL * NL * b;
    if (f < 4) return a + NL * side + NL * NL * b;
    return a + NL * b + NL * NL * side;
}

// Extended-block position of own point p = i + NH j + NH^2 k.
template <int NH>
__device__ __forceinline__ int own_ext(int p)
{
    constexpr int NL = NH + 2;
    const int i = p % NH, j = (p / NH) % NH, k = p / (NH * NH);
    return (i + 1) + NL * (j + 1) + NL * NL * (k + 1);
}

// Local solves of one Schwarz level (h1mg_schwarz_part1, src/hsmg.f:1052-1079):
//   mask r in place (the reference's side effect, :1067); build the
//   extended element (own points, each ghost face = the neighbour's point
//   one inside, delivered as (own + neighbour) - own like the (n+2)^3
//   exchange of :1071-1076; edges and corners zero); e = S D^-1 S^T r with
//   S = S_z x S_y x S_x (h1mg_do_fast, :1297-1320): contractions along x,
//   y, z with the S^T factors, the diagonal (l_x + l_y + l_z, cut to zero
//   at or below the element's eps), then x, y, z with the S factors; write
//   the own part and the ghost faces of e.
// The element's block sits in LDS; each contraction reads a thread's line
// into registers and writes it back in place.  T is the arithmetic of the
// local solve (FP64, or FP32 for the single-precision smoother); the level
// vectors stay FP64.
template <typename T, int NH>
__global__ void __launch_bounds__(FdmShape<NH>::NT)
    k_fdm(FdmArgs a, const T *__restrict__ ops, const T *__restrict__ epsv)
{
    using S = FdmShape<NH>;
    constexpr int NL = S::NL, L2 = S::L2, L3 = S::L3, EPB = S::EPB, OPS = S::OPS;
    constexpr int H2 = NH * NH, N3 = H2 * NH, G2 = 6 * H2;
    __shared__ T blk[EPB][L3];
    __shared__ T op[EPB][3][OPS];
    const int e0 = blockIdx.x * EPB, tid = threadIdx.x, nelt = a.nelt;
    {
        constexpr int IOP = (EPB * 3 * OPS + S::NT - 1) / S::NT;
        T vop[IOP];
#pragma unroll
        for (int it = 0; it < IOP; ++it) {
            const int q = tid + it * S::NT, el = q / (3 * OPS), w = q - el * (3 * OPS),
                      d = w / OPS;
            vop[it] = T(0);
            if (q < EPB * 3 * OPS && e0 + el < nelt)
                vop[it] = ops[int64_t(a.eop[3 * (e0 + el) + d]) * OPS + (w - d * OPS)];
        }
#pragma unroll
        for (int it = 0; it < IOP; ++it) {
            const int q = tid + it * S::NT;
            if (q < EPB * 3 * OPS) (&op[0][0][0])[q] = vop[it];
        }
    }
    // edges and corners of the extended block stay zero
    for (int q = tid; q < EPB * L3; q += S::NT) (&blk[0][0])[q] = T(0);
    __syncthreads();
    // Gathers first (unrolled, no stores in between, so every thread's
    // loads are in flight together), then the LDS writes.  Dirichlet
    // points read as zero here; r itself is masked at the end.
    constexpr int IO = (EPB * N3 + S::NT - 1) / S::NT, IG = (EPB * G2 + S::NT - 1) / S::NT;
    double vown[IO], vgh[IG];
#pragma unroll
    for (int it = 0; it < IO; ++it) {
        const int q = tid + it * S::NT, el = q / N3, e = e0 + el;
        vown[it] = 0.0;
        if (q < EPB * N3 && e < nelt) {
            const int64_t m = int64_t(e) * N3 + a.pos[q - el * N3];
            vown[it] = a.bnd[m] ? 0.0 : a.r[m];
        }
    }
#pragma unroll
    for (int it = 0; it < IG; ++it) {
        const int q = tid + it * S::NT, el = q / G2, e = e0 + el;
        vgh[it] = 0.0;
        if (q < EPB * G2 && e < nelt) {
            const int64_t slot = int64_t(e) * G2 + (q - el * G2);
            const int32_t sn = a.gsrc[slot];
            if (sn >= 0) {
                const int32_t o = a.gin[slot];
                const double vo = a.bnd[o] ? 0.0 : a.r[o];
                const double vs = a.bnd[sn] ? 0.0 : a.r[sn];
                vgh[it] = (vo + vs) - vo;
            }
        }
    }
#pragma unroll
    for (int it = 0; it < IO; ++it) {
        const int q = tid + it * S::NT, el = q / N3;
        if (q < EPB * N3 && e0 + el < nelt) blk[el][own_ext<NH>(q - el * N3)] = T(vown[it]);
    }
#pragma unroll
    for (int it = 0; it < IG; ++it) {
        const int q = tid + it * S::NT, el = q / G2;
        if (q < EPB * G2 && e0 + el < nelt) {
            const int w = q - el * G2;
            blk[el][ghost_ext<NH>(w / H2, w - (w / H2) * H2)] = T(vgh[it]);
        }
    }
    __syncthreads();
    const int el = tid / L2, l = tid - el * L2, la = l % NL, lb = l / NL;
    const bool act = el < EPB && e0 + el < nelt;
    T *b = blk[act ? el : 0];
    const T *ox = op[act ? el : 0][0], *oy = op[act ? el : 0][1], *oz = op[act ? el : 0][2];
    T v[NL];
    // S^T r: x lines (la, lb) = (j, k), y lines (i, k), z lines (i, j)
    if (act) {
        line_load(v, b, NL * l, 1);
        line_apply<T, NL>(v, ox + L2);
        line_store(v, b, NL * l, 1);
    }
    __syncthreads();
    if (act) {
        line_load(v, b, la + L2 * lb, NL);
        line_apply<T, NL>(v, oy + L2);
        line_store(v, b, la + L2 * lb, NL);
    }
    __syncthreads();
    if (act) {
        line_load(v, b, l, L2);
        line_apply<T, NL>(v, oz + L2);
        const T eps = epsv[e0 + el], lxy = ox[2 * L2 + la] + oy[2 * L2 + lb];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const T diag = lxy + oz[2 * L2 + k];
            const T d = (diag > eps) ? T(1) / diag : T(0);
            v[k] = d * v[k];
        }
        line_store(v, b, l, L2);
    }
    __syncthreads();
    if (act) {
        line_load(v, b, NL * l, 1);
        line_apply<T, NL>(v, ox);
        line_store(v, b, NL * l, 1);
    }
    __syncthreads();
    if (act) {
        line_load(v, b, la + L2 * lb, NL);
        line_apply<T, NL>(v, oy);
        line_store(v, b, la + L2 * lb, NL);
    }
    __syncthreads();
    if (act) {
        line_load(v, b, l, L2);
        line_apply<T, NL>(v, oz);
        line_store(v, b, l, L2);
    }
    __syncthreads();
    {
        int64_t mo[IO];
        uint8_t bo[IO];
#pragma unroll
        for (int it = 0; it < IO; ++it) {
            const int q = tid + it * S::NT, el2 = q / N3, e = e0 + el2;
            mo[it] = -1;
            bo[it] = 0;
            if (q < EPB * N3 && e < nelt) {
                mo[it] = int64_t(e) * N3 + a.pos[q - el2 * N3];
                bo[it] = a.bnd[mo[it]];
            }
        }
#pragma unroll
        for (int it = 0; it < IO; ++it) {
            const int q = tid + it * S::NT, el2 = q / N3;
            if (mo[it] < 0) continue;
            a.eown[mo[it]] = double(blk[el2][own_ext<NH>(q - el2 * N3)]);
            if (bo[it]) a.r[mo[it]] = 0.0;  // the caller's r leaves masked (src/hsmg.f:1067)
        }
    }
    for (int q = tid; q < EPB * G2; q += S::NT) {
        const int el2 = q / G2, e = e0 + el2;
        if (e >= nelt) continue;
        const int w = q - el2 * G2;
        a.ghost[int64_t(e) * G2 + w] = double(blk[el2][ghost_ext<NH>(w / H2, w - (w / H2) * H2)]);
    }
}

template <typename T, int NH>
int launch_fdm_nh(const FdmArgs &a, const T *ops, const T *eps)
{
    using S = FdmShape<NH>;
    hipLaunchKernelGGL((k_fdm<T, NH>), dim3((a.nelt + S::EPB - 1) / S::EPB), dim3(S::NT), 0,
                       stream(), a, ops, eps);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

template <typename T>
int launch_fdm(const FdmArgs &a, const T *ops, const T *eps, int nh)
{
    switch (nh) {
    case 3: return launch_fdm_nh<T, 3>(a, ops, eps);
    case 4: return launch_fdm_nh<T, 4>(a, ops, eps);
    case 5: return launch_fdm_nh<T, 5>(a, ops, eps);
    case 6: return launch_fdm_nh<T, 6>(a, ops, eps);
    case 7: return launch_fdm_nh<T, 7>(a, ops, eps);
    case 8: return launch_fdm_nh<T, 8>(a, ops, eps);
    case 9: return launch_fdm_nh<T, 9>(a, ops, eps);
    case 10: return launch_fdm_nh<T, 10>(a, ops, eps);
    case 11: return launch_fdm_nh<T, 11>(a, ops, eps);
    case 12: return launch_fdm_nh<T, 12>(a, ops, eps);
    case 13: return launch_fdm_nh<T, 13>(a, ops, eps);
    case 14: return launch_fdm_nh<T, 14>(a, ops, eps);
    case 15: return launch_fdm_nh<T, 15>(a, ops, eps);
    case 16: return launch_fdm_nh<T, 16>(a, ops, eps);
    default: return fail(NEKBONE_ERR_UNSUPPORTED, "hsmg: level size outside 3..16");
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
    const FdmArgs a{r, L.d_bnd, L.d_pos, L.d_gsrc, L.d_gin, L.d_eop, L.d_eown, L.d_ghost, P->nelt};
    if (int rc = P->hsmg_fp32 ? launch_fdm<float>(a, L.d_ops32, L.d_eps32, L.nh)
                              : launch_fdm<double>(a, L.d_ops, L.d_eps, L.nh))
        return rc;
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
    // coarse level: e1 = mask(QQ^T mask(J^T r2)) (hsmg_coarse_solve, :945-957),
    // or with the direct coarse solve e1 = Q A_c^-1 Q^T mask(J^T r2)
    if (int rc = rstr(P, L[1], L[2], L[1].d_e, rcur, true)) return rc;
    if (int rc = launch_gs_map(L[1].gs, L[1].d_e)) return rc;
    if (P->hsmg_coarse)
        if (int rc = hsmg_coarse_apply(P, L[1].d_e)) return rc;
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

extern "C" int nekbone_set_hsmg_fp32(nekbone_problem *P, int on)
{
    if (int rc = check(P)) return rc;
    P->hsmg_fp32 = on ? 1 : 0;
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
