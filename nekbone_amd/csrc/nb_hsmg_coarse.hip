// Direct FP64 coarse solve for the HSMG preconditioner (optional; the
// reference's coarse level is hsmg_coarse_solve, src/hsmg.f:945-957,
// which only sums the restricted residual, e = QQ^T r).
//
// The coarse space is the order-1 level of the hierarchy: one trilinear
// function per mesh vertex, reached from the fine level through the
// hierarchy's interpolations, whose product is the trilinear
// interpolation J (8 vertex values -> NX^3 GLL points of an element).
// Its operator is the Galerkin projection of the fine one,
//   A_c = P^T (Q^T A_L Q + 0.1 I) P = sum_e J^T A_L^e J + 0.1 sum_e J^T C_e J
// on the unmasked mesh vertices (C_e = inverse multiplicity of the
// element's points), which in the numbers of the N=7 deformed test mesh
// takes a third fewer CG iterations than re-discretising at N = 1.
//
// Set-up (once): the eight hat functions J e_q are spread over every
// element on the device, the fine element kernel applies A_L to each, and
// a reduction kernel forms the 8x8 element matrices; the host assembles
// the dense A_c, which rocSOLVER factors and inverts (potrf, potri); the
// lower triangle is mirrored.  Apply, replacing the coarse step of
// h1mg_solve: after the level's dssum every copy of a vertex holds the
// summed residual, so
//   r_g[j] = e[prim[j]],  x_g = A_c^-1 r_g  (one workgroup per row, rows of
//   the symmetric inverse read as contiguous columns; the inverse, 91 MB for
//   the 15^3 interior vertices of a 16^3 mesh, is read once per apply),
//   e[q] = x_g[cidx[q]] or 0 on the mask.
// Every sum has a fixed order, so results are reproducible run to run.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <map>
#include <vector>

#include "nb_device.h"

namespace nb {
namespace {

// u = J e_q on every element: point p (reference order) of element e, at
// e*n3 + perm[p] in device order, gets the hat value jq[p].
__global__ void k_fill_hat(double *__restrict__ u, const double *__restrict__ jq,
                           const int32_t *__restrict__ perm, int n3, int64_t n)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t e = t / n3;
    const int p = int(t - e * n3);
    u[e * n3 + perm[p]] = jq[p];
}

// Column q of the element matrices: ae[(e*8 + a)*8 + q] =
// sum_p J(p,a) (w[e,p] + 0.1 c[e,p] J(p,q)), p ascending.  j is NX^3 x 8
// (j[p*8 + a]); w = A_L J e_q and c in device order.
__global__ void k_galerkin_col(double *__restrict__ ae, const double *__restrict__ w,
                               const double *__restrict__ c, const double *__restrict__ j,
                               const int32_t *__restrict__ perm, int n3, int q, int64_t nelt)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nelt * 8) return;
    const int64_t e = t / 8;
    const int a = int(t - e * 8);
    double s = 0.0;
    for (int p = 0; p < n3; ++p) {
        const int64_t d = e * n3 + perm[p];
        s = s + j[p * 8 + a] * (w[d] + 0.1 * c[d] * j[p * 8 + q]);
    }
    ae[(e * 8 + a) * 8 + q] = s;
}

__global__ void k_mirror_lower(double *__restrict__ a, int n)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(n) * n) return;
    const int i = int(t % n), j = int(t / n);  // column-major (i, j)
    if (i < j) a[t] = a[j + int64_t(n) * i];
}

// r_g[j] = e[prim[j]]: the summed residual of each unknown.
__global__ void k_coarse_gather(double *__restrict__ rg, const double *__restrict__ e,
                                const int32_t *__restrict__ prim, int nc,
                                const CgScalars *__restrict__ guard)
{
    if (guard && guard->done) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nc) rg[j] = e[prim[j]];
}

// x_g = A^-1 r_g, one 256-thread workgroup per row (a contiguous column
// of the symmetric inverse): every lane's loads of the row (up to
// kGemvU) in flight together, its partial sums added in a fixed order,
// then the workgroup's fixed tree.  (One wave per row with four or sixteen
// loads in flight per lane left the ~3 waves per SIMD that 3,375 rows
// give latency-bound at 3.7 TB/s or below; this form reaches 4.0 TB/s,
// 23 us at 16^3 N=7.)
constexpr int kGemvThreads = 256;
constexpr int kGemvU = 16;  // row length up to kGemvThreads * kGemvU per pass
__global__ void __launch_bounds__(kGemvThreads)
    k_coarse_gemv(double *__restrict__ xg, const double *__restrict__ ainv,
                  const double *__restrict__ rg, int nc, const CgScalars *__restrict__ guard)
{
    if (guard && guard->done) return;
    const int row = blockIdx.x;
    const double *col = ainv + int64_t(row) * nc;
    double t = 0.0;
    for (int i0 = threadIdx.x; i0 < nc; i0 += kGemvThreads * kGemvU) {
        double v[kGemvU];
#pragma unroll
        for (int u = 0; u < kGemvU; ++u) {
            const int i = i0 + kGemvThreads * u;
            v[u] = i < nc ? col[i] * rg[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kGemvU; ++u) t = t + v[u];
    }
    double dummy = 0.0;
    block_sum2<kGemvThreads>(t, dummy);
    if (threadIdx.x == 0) xg[row] = t;
}

__global__ void k_coarse_scatter(double *__restrict__ e, const double *__restrict__ xg,
                                 const int32_t *__restrict__ cidx, int64_t n,
                                 const CgScalars *__restrict__ guard)
{
    if (guard && guard->done) return;
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int32_t j = cidx[q];
    e[q] = j >= 0 ? xg[j] : 0.0;
}

}  // namespace

void hsmg_coarse_free(Hsmg *M)
{
    dfree(M->d_ainv);
    dfree(M->d_cprim);
    dfree(M->d_cidx);
    dfree(M->d_xc);
    dfree(M->d_rc);
    M->nc = 0;
}

int hsmg_coarse_build(nekbone_problem *P)
{
    Hsmg *M = P->mg;
    if (M->d_ainv) return NEKBONE_OK;
    const int nelt = P->nelt, nx = P->nx1;
    const int64_t n1 = int64_t(8) * nelt;  // order-1 level points

    // unknowns: global vertex ids of points not on the level's mask
    std::vector<int64_t> glo;
    box_numbering(glo, 2, P->nelx, P->nely, P->nelz, P->lglel);
    std::vector<uint8_t> bnd(static_cast<size_t>(n1));
    if (int rc = d2h(bnd.data(), M->L[1].d_bnd, size_t(n1))) return rc;
    std::map<int64_t, int32_t> id;
    std::vector<int32_t> cidx(static_cast<size_t>(n1), -1), prim;
    for (int64_t q = 0; q < n1; ++q) {
        if (bnd[size_t(q)]) continue;
        auto it = id.find(glo[size_t(q)]);
        if (it == id.end()) {
            it = id.emplace(glo[size_t(q)], int32_t(prim.size())).first;
            prim.push_back(int32_t(q));
        }
        cidx[size_t(q)] = it->second;
    }
    const int nc = int(prim.size());
    if (nc == 0) return fail(NEKBONE_ERR_UNSUPPORTED, "hsmg coarse: no unmasked vertices");
    if (nc > kMaxCoarse)
        return fail(NEKBONE_ERR_UNSUPPORTED, "hsmg coarse: more than 8192 coarse unknowns");

    // element matrices J^T (A_L + 0.1 C_e) J, one hat function at a time
    const int n3 = nx * nx * nx;
    std::vector<double> jh(size_t(n3) * 8);  // J(p, a) at jh[p*8 + a]
    for (int k = 0; k < nx; ++k)
        for (int jj = 0; jj < nx; ++jj)
            for (int i = 0; i < nx; ++i) {
                const int p = i + nx * jj + nx * nx * k;
                const double zi = P->zgll[i], zj = P->zgll[jj], zk = P->zgll[k];
                for (int a = 0; a < 8; ++a) {
                    const double li = (a & 1) ? 0.5 * (1.0 + zi) : 0.5 * (1.0 - zi);
                    const double lj = (a & 2) ? 0.5 * (1.0 + zj) : 0.5 * (1.0 - zj);
                    const double lk = (a & 4) ? 0.5 * (1.0 + zk) : 0.5 * (1.0 - zk);
                    jh[size_t(p) * 8 + size_t(a)] = li * lj * lk;
                }
            }
    std::vector<double> jcol(static_cast<size_t>(n3));
    double *d_j = nullptr, *d_jq = nullptr, *d_ae = nullptr;
    int rc = dalloc(&d_j, int64_t(n3) * 8);
    if (!rc) rc = dalloc(&d_jq, int64_t(n3));
    if (!rc) rc = dalloc(&d_ae, int64_t(nelt) * 64);
    hipError_t err = hipSuccess;
    if (!rc) rc = h2d(d_j, jh.data(), sizeof(double) * jh.size());
    for (int q = 0; q < 8 && !rc && err == hipSuccess; ++q) {
        for (int p = 0; p < n3; ++p) jcol[size_t(p)] = jh[size_t(p) * 8 + size_t(q)];
        err = hipMemcpyAsync(d_jq, jcol.data(), sizeof(double) * jcol.size(),
                             hipMemcpyHostToDevice, stream());
        if (err == hipSuccess) err = hipStreamSynchronize(stream());
        if (err != hipSuccess) break;
        hipLaunchKernelGGL(k_fill_hat, dim3(unsigned((P->n + 255) / 256)), dim3(256), 0, stream(),
                           P->d_t, static_cast<const double *>(d_jq),
                           static_cast<const int32_t *>(P->d_perm), n3, P->n);
        rc = launch_ax_local(P, P->d_t2, P->d_t);
        if (rc) break;
        hipLaunchKernelGGL(k_galerkin_col, dim3(unsigned((int64_t(nelt) * 8 + 255) / 256)),
                           dim3(256), 0, stream(), d_ae, static_cast<const double *>(P->d_t2),
                           static_cast<const double *>(P->d_c),
                           static_cast<const double *>(d_j),
                           static_cast<const int32_t *>(P->d_perm), n3, q, int64_t(nelt));
        err = hipGetLastError();
    }
    std::vector<double> ae(size_t(nelt) * 64);
    if (!rc && err == hipSuccess)
        err = hipMemcpyAsync(ae.data(), d_ae, sizeof(double) * ae.size(), hipMemcpyDeviceToHost,
                             stream());
    if (!rc && err == hipSuccess) err = hipStreamSynchronize(stream());
    dfree(d_j);
    dfree(d_jq);
    dfree(d_ae);
    if (rc) return rc;
    NB_HIP(err);

    std::vector<double> A(size_t(nc) * nc, 0.0);
    for (int e = 0; e < nelt; ++e)
        for (int b = 0; b < 8; ++b) {
            const int32_t jc = cidx[size_t(int64_t(e) * 8 + b)];
            if (jc < 0) continue;
            for (int a = 0; a < 8; ++a) {
                const int32_t ir = cidx[size_t(int64_t(e) * 8 + a)];
                if (ir >= 0)
                    A[size_t(ir) + size_t(nc) * size_t(jc)] += ae[size_t((int64_t(e) * 8 + a) * 8 + b)];
            }
        }

    if (int rc = dalloc(&M->d_ainv, int64_t(nc) * nc)) return rc;
    if (int rc = h2d(M->d_ainv, A.data(), sizeof(double) * A.size())) return rc;
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success)
        return fail(NEKBONE_ERR_HIP, "hsmg coarse: rocblas handle");
    rocblas_set_stream(h, stream());
    rocblas_int *d_info = nullptr;
    int info = 0;
    rocblas_status st = rocblas_status_success;
    if (dalloc(&d_info, 1) != NEKBONE_OK) st = rocblas_status_memory_error;
    if (st == rocblas_status_success)
        st = rocsolver_dpotrf(h, rocblas_fill_lower, nc, M->d_ainv, nc, d_info);
    if (st == rocblas_status_success &&
        hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, stream()) == hipSuccess &&
        hipStreamSynchronize(stream()) == hipSuccess && info == 0)
        st = rocsolver_dpotri(h, rocblas_fill_lower, nc, M->d_ainv, nc, d_info);
    else if (st == rocblas_status_success)
        st = rocblas_status_internal_error;
    if (st == rocblas_status_success &&
        (hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, stream()) !=
             hipSuccess ||
         hipStreamSynchronize(stream()) != hipSuccess || info != 0))
        st = rocblas_status_internal_error;
    dfree(d_info);
    rocblas_destroy_handle(h);
    if (st != rocblas_status_success)
        return fail(NEKBONE_ERR_HIP, "hsmg coarse: Cholesky inverse of the coarse operator failed");
    hipLaunchKernelGGL(k_mirror_lower, dim3(unsigned((int64_t(nc) * nc + 255) / 256)), dim3(256),
                       0, stream(), M->d_ainv, nc);
    NB_HIP(hipGetLastError());

    if (int rc = dalloc(&M->d_cprim, int64_t(nc))) return rc;
    if (int rc = dalloc(&M->d_cidx, n1)) return rc;
    if (int rc = dalloc(&M->d_xc, int64_t(nc))) return rc;
    if (int rc = dalloc(&M->d_rc, int64_t(nc))) return rc;
    if (int rc = h2d(M->d_cprim, prim.data(), sizeof(int32_t) * prim.size())) return rc;
    if (int rc = h2d(M->d_cidx, cidx.data(), sizeof(int32_t) * cidx.size())) return rc;
    M->nc = nc;
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

// e <- Q A_c^-1 Q^T e on the order-1 level (e already summed by the
// level's dssum: every copy holds Q^T e of its vertex).  guard (may be
// null): the kernels return at once once the solve has converged.
int hsmg_coarse_apply(nekbone_problem *P, double *e, const CgScalars *guard)
{
    Hsmg *M = P->mg;
    const int nc = M->nc;
    hipLaunchKernelGGL(k_coarse_gather, dim3(unsigned((nc + 255) / 256)), dim3(256), 0, stream(),
                       M->d_rc, static_cast<const double *>(e),
                       static_cast<const int32_t *>(M->d_cprim), nc, guard);
    hipLaunchKernelGGL(k_coarse_gemv, dim3(unsigned(nc)), dim3(kGemvThreads),
                       0, stream(), M->d_xc, static_cast<const double *>(M->d_ainv),
                       static_cast<const double *>(M->d_rc), nc, guard);
    NB_HIP(hipGetLastError());
    const int64_t n1 = M->L[1].n;
    hipLaunchKernelGGL(k_coarse_scatter, dim3(unsigned((n1 + 255) / 256)), dim3(256), 0, stream(),
                       e, static_cast<const double *>(M->d_xc),
                       static_cast<const int32_t *>(M->d_cidx), n1, guard);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

}  // namespace nb

using namespace nb;

extern "C" int nekbone_set_hsmg_coarse(nekbone_problem *P, int mode)
{
    if (int rc = check(P)) return rc;
    if (mode != 0 && mode != 1) return fail(NEKBONE_ERR_ARG, "hsmg coarse: mode 0 or 1");
    if (mode == 1) {
        if (!P->mg) return fail(NEKBONE_ERR_STATE, "hsmg coarse: nekbone_setup_hsmg not called");
        // the coarse buffers are rebuilt: a graph captured over the old ones
        // must not replay
        drop_problem_graph(P);
        if (int rc = hsmg_coarse_build(P)) {
            hsmg_coarse_free(P->mg);
            return rc;
        }
    }
    P->hsmg_coarse = mode;
    return NEKBONE_OK;
}
