// Reference-named, Fortran-callable entry points (trailing underscore, all
// arguments by reference).  They act on the problem bound by
// nekbone_setup_, standing in for the COMMON state the reference reads
// (/dimn/, /dxyz/, /mymask/, /comm_handles/).  Host arrays in and out;
// all arithmetic runs on the device through the same kernels as the
// explicit API.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>

#include "nb_internal.h"

namespace nb {
namespace {

struct GsHandle {
    int64_t n = 0;
    std::vector<int64_t> glo;
    std::unique_ptr<nekbone_problem, int (*)(nekbone_problem *)> host{nullptr,
                                                                       nekbone_problem_destroy};
    // groups in the jl order, CSR over reference local indices
    struct {
        int64_t ngroups = 0, nshared = 0;
        int32_t *d_off = nullptr, *d_idx = nullptr;
    } map;
    double *d_u = nullptr;
};

// Content fingerprint of a host array: a 64-bit multiply-xorshift hash of
// every word in four interleaved lanes (one streaming pass).  The legacy
// entry points take g and c on every call, as the reference does; they
// upload them again only when the address or the contents changed, so an
// in-place update by the caller is not missed.
uint64_t fingerprint(const double *a, int64_t count)
{
    uint64_t h[4] = {0x243F6A8885A308D3ull, 0x13198A2E03707344ull, 0xA4093822299F31D0ull,
                     0x082EFA98EC4E6C89ull};
    const uint64_t *w = reinterpret_cast<const uint64_t *>(a);
    int64_t i = 0;
    for (; i + 4 <= count; i += 4)
        for (int l = 0; l < 4; ++l) {
            uint64_t v = (h[l] ^ w[i + l]) * 0x9E3779B97F4A7C15ull;
            h[l] = v ^ (v >> 29);
        }
    for (; i < count; ++i) {
        uint64_t v = (h[0] ^ w[i]) * 0x9E3779B97F4A7C15ull;
        h[0] = v ^ (v >> 29);
    }
    uint64_t out = uint64_t(count);
    for (int l = 0; l < 4; ++l) {
        uint64_t v = (out ^ h[l]) * 0xBF58476D1CE4E5B9ull;
        out = v ^ (v >> 31);
    }
    return out;
}

struct LegacyState {
    nekbone_problem *P = nullptr;
    const double *g_ptr = nullptr;  // host g last uploaded, and its fingerprint
    const double *c_ptr = nullptr;  // host c last uploaded, and its fingerprint
    uint64_t g_fp = 0, c_fp = 0;
    int precond = NEKBONE_PRECOND_IDENTITY;
    bool jacobi_on = false;  // nekbone_use_jacobi_: kept across meshes
    std::map<int, GsHandle> gs;
    int next_handle = 1;
    nekbone_problem *elem = nullptr;  // one-element problem for ax_e_
    double *d_scratch = nullptr;
    int64_t scratch_len = 0;
};

LegacyState &S()
{
    static LegacyState s;
    return s;
}

void report(int rc, const char *where)
{
    if (rc != NEKBONE_OK)
        std::fprintf(stderr, "nekbone_amd: %s failed (%d): %s\n", where, rc, nekbone_last_error());
}

int scratch(int64_t count, double **out)
{
    LegacyState &L = S();
    if (L.scratch_len < count) {
        dfree(L.d_scratch);
        L.scratch_len = 0;
        if (int rc = dalloc(&L.d_scratch, count)) return rc;
        L.scratch_len = count;
    }
    *out = L.d_scratch;
    return NEKBONE_OK;
}

__global__ void k_gs_plain(double *__restrict__ u, const int32_t *__restrict__ off,
                           const int32_t *__restrict__ idx, int64_t ngroups)
{
    const int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= ngroups) return;
    const int a = off[gid], b = off[gid + 1];
    double t = u[idx[a]];
    for (int q = a + 1; q < b; ++q) t += u[idx[q]];
    for (int q = a; q < b; ++q) u[idx[q]] = t;
}

// local_grad3 (src/cg.f:170-187) for one element, one thread per point,
// sums in mxf order.
__global__ void k_grad3(double *ur, double *us, double *ut, const double *u, const double *D,
                        const double *Dt, int nx)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int nxyz = nx * nx * nx;
    if (p >= nxyz) return;
    const int i = p % nx, j = (p / nx) % nx, k = p / (nx * nx);
    double a = D[i] * u[0 + nx * j + nx * nx * k];
    double b = u[i + nx * 0 + nx * nx * k] * Dt[0 + nx * j];
    double c = u[i + nx * j] * Dt[0 + nx * k];
    for (int m = 1; m < nx; ++m) {
        a = a + D[i + nx * m] * u[m + nx * j + nx * nx * k];
        b = b + u[i + nx * m + nx * nx * k] * Dt[m + nx * j];
        c = c + u[i + nx * j + nx * nx * m] * Dt[m + nx * k];
    }
    ur[p] = a;
    us[p] = b;
    ut[p] = c;
}

// local_grad3_t (src/cg.f:189-212): u = (Dt ur + us D) + ut D.
__global__ void k_grad3_t(double *out, const double *ur, const double *us, const double *ut,
                          const double *D, const double *Dt, int nx)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int nxyz = nx * nx * nx;
    if (p >= nxyz) return;
    const int i = p % nx, j = (p / nx) % nx, k = p / (nx * nx);
    double a = Dt[i] * ur[0 + nx * j + nx * nx * k];
    double b = us[i + nx * 0 + nx * nx * k] * D[0 + nx * j];
    double c = ut[i + nx * j] * D[0 + nx * k];
    for (int m = 1; m < nx; ++m) {
        a = a + Dt[i + nx * m] * ur[m + nx * j + nx * nx * k];
        b = b + us[i + nx * m + nx * nx * k] * D[m + nx * j];
        c = c + ut[i + nx * j + nx * nx * m] * D[m + nx * k];
    }
    out[p] = (a + b) + c;
}

__global__ void k_copy(double *dst, const double *src, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_mul(double *dst, const double *a, const double *b, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = a[i] * b[i];
}

// Upload g and c when the caller hands a different array than last time,
// or the same array with different contents.
int sync_geometry(const double *g, const double *c)
{
    LegacyState &L = S();
    if (g) {
        const uint64_t fp = fingerprint(g, 6 * L.P->n);
        if (g != L.g_ptr || fp != L.g_fp) {
            if (int rc = nekbone_set_geometry(L.P, g)) return rc;
            L.g_ptr = g;
            L.g_fp = fp;
            if (L.precond == NEKBONE_PRECOND_JACOBI)
                if (int rc = nekbone_setup_jacobi(L.P)) return rc;
        }
    }
    if (c) {
        const uint64_t fp = fingerprint(c, L.P->n);
        if (c != L.c_ptr || fp != L.c_fp) {
            if (int rc = set_c_host(L.P, c)) return rc;
            L.c_ptr = c;
            L.c_fp = fp;
        }
    }
    return NEKBONE_OK;
}

// The reference's residual line, format ('cg:',i4,1p4e12.4)
// (src/cg.f:43,76); nvals of rnorm, alpha, beta, pap.
void cg_line(int iter, const double *v, int nvals)
{
    std::printf("cg:%4d", iter);
    for (int q = 0; q < nvals; ++q) std::printf("%12.4E", v[q]);
    std::printf("\n");
    std::fflush(stdout);
}

int bound(const char *where)
{
    if (!S().P) return fail(NEKBONE_ERR_STATE, std::string(where) + ": nekbone_setup_ not called");
    return NEKBONE_OK;
}

}  // namespace
}  // namespace nb

using namespace nb;

// COMMON /cflops/ flop_a, flop_cg, mflops (src/INPUT): updated when the
// Fortran host defines it, as src/cg.f:81,122 do.
extern "C" {
struct nb_cflops {
    double flop_a, flop_cg, mflops;
};
extern nb_cflops cflops_ __attribute__((weak));
}

extern "C" void nekbone_setup_(const int *nelt, const int *nx1, const int *nelx, const int *nely,
                               const int *nelz, const int32_t *lglel, const double *dxm1,
                               const int *gsh)
{
    LegacyState &L = S();
    auto it = L.gs.find(*gsh);
    if (it == L.gs.end()) {
        report(fail(NEKBONE_ERR_ARG, "nekbone_setup_: unknown gs handle"), "nekbone_setup_");
        return;
    }
    if (L.P) nekbone_problem_destroy(L.P);
    L.P = nullptr;
    L.g_ptr = L.c_ptr = nullptr;
    L.g_fp = L.c_fp = 0;
    int rc = nekbone_problem_create(&L.P, *nelt, *nx1, *nelx, *nely, *nelz, lglel, dxm1,
                                    it->second.glo.data(), nullptr, nullptr);
    // a new mesh starts with the solveM of the build: identity, or Jacobi
    // when selected; HSMG again only after h1mg_setup_ for this mesh, as
    // src/driver.f:46 calls it per mesh
    L.precond = L.jacobi_on ? NEKBONE_PRECOND_JACOBI : NEKBONE_PRECOND_IDENTITY;
    if (!rc && L.jacobi_on) rc = nekbone_setup_jacobi(L.P);
    report(rc, "nekbone_setup_");
}

extern "C" void nekbone_use_jacobi_(const int *on)
{
    LegacyState &L = S();
    L.jacobi_on = *on != 0;
    L.precond = (*on) ? NEKBONE_PRECOND_JACOBI : NEKBONE_PRECOND_IDENTITY;
    if (L.P && L.precond == NEKBONE_PRECOND_JACOBI)
        report(nekbone_setup_jacobi(L.P), "nekbone_use_jacobi_");
}

extern "C" void cg_(double *x, const double *f, const double *g, const double *c, double *r,
                    double *w, double *p, double *z, const int *n, const int *niter,
                    double *flop_cg)
{
    LegacyState &L = S();
    int rc = bound("cg_");
    if (!rc && *n != L.P->n) rc = fail(NEKBONE_ERR_ARG, "cg_: n differs from the bound mesh");
    if (!rc) rc = sync_geometry(g, c);
    std::vector<double> hist(size_t(1 + 3 * std::max(*niter, 0)));
    if (!rc) rc = nekbone_cg(L.P, x, f, r, *niter, L.precond, hist.data());
    if (!rc) {
        // src/cg.f:43 and :76: the initial residual norm, then after the
        // loop iter = niter+1 with the last rnorm, alpha, beta and pap
        const double r0 = std::sqrt(hist[0]);
        cg_line(0, &r0, 1);
        const int m = *niter;
        double v[4] = {r0, 0.0, 0.0, 0.0};
        if (m >= 1) {
            const double rtz1 = hist[size_t(3 * m - 2)], pap = hist[size_t(3 * m - 1)];
            v[0] = std::sqrt(hist[size_t(3 * m)]);
            v[1] = rtz1 / pap;
            v[2] = m == 1 ? 0.0 : rtz1 / hist[size_t(3 * m - 5)];
            v[3] = pap;
        }
        cg_line(m + 1, v, 4);
        if (w) rc = download_ref(L.P, w, L.P->d_w);
        if (!rc && p) rc = download_ref(L.P, p, L.P->d_p);
        (void)z;  // z is scratch in the reference; the fused path never stores it
        // src/cg.f:79-81: after the loop iter = niter+1
        const double iter = double(*niter + 1);
        if (flop_cg) *flop_cg += iter * 15.0 * double(*n);
        const double nxyz = double(L.P->nx1) * L.P->nx1 * L.P->nx1;
        if (&cflops_)
            cflops_.flop_a += double(*niter) * (19.0 * nxyz + 12.0 * L.P->nx1 * nxyz) * L.P->nelt;
    }
    report(rc, "cg_");
}

extern "C" void solvem_(double *z, double *r, const int *n)
{
    LegacyState &L = S();
    int rc = bound("solvem_");
    if (!rc && *n != L.P->n) rc = fail(NEKBONE_ERR_ARG, "solvem_: n differs from the bound mesh");
    if (rc) return report(rc, "solvem_");
    if (L.precond == NEKBONE_PRECOND_HSMG)  // MGRID build: h1mg_solve (src/cg.f:91)
        return report(nekbone_hsmg_solve(L.P, z, r), "solvem_");
    const int64_t nn = L.P->n;
    if (int rc2 = upload_dev(L.P, L.P->d_t2, r)) return report(rc2, "solvem_");
    if (L.precond == NEKBONE_PRECOND_JACOBI && L.P->d_dinv)
        hipLaunchKernelGGL(k_mul, dim3(unsigned((nn + 255) / 256)), dim3(256), 0, stream(),
                           L.P->d_w, L.P->d_t2, L.P->d_dinv, nn);
    else
        hipLaunchKernelGGL(k_copy, dim3(unsigned((nn + 255) / 256)), dim3(256), 0, stream(),
                           L.P->d_w, L.P->d_t2, nn);
    report(download_ref(L.P, z, L.P->d_w), "solvem_");
}

extern "C" void h1mg_setup_(void)
{
    LegacyState &L = S();
    int rc = bound("h1mg_setup_");
    if (!rc) rc = nekbone_setup_hsmg(L.P);
    if (!rc) L.precond = NEKBONE_PRECOND_HSMG;
    report(rc, "h1mg_setup_");
}

extern "C" void h1mg_solve_(double *z, double *rhs, const int *n)
{
    LegacyState &L = S();
    int rc = bound("h1mg_solve_");
    if (!rc && *n != L.P->n) rc = fail(NEKBONE_ERR_ARG, "h1mg_solve_: n differs from the bound mesh");
    if (!rc) rc = nekbone_hsmg_solve(L.P, z, rhs);
    report(rc, "h1mg_solve_");
}

extern "C" void ax_(double *w, const double *u, const double *gxyz, double *ur, double *us,
                    double *ut, double *wk, const int *n)
{
    (void)ur, (void)us, (void)ut, (void)wk;
    LegacyState &L = S();
    int rc = bound("ax_");
    if (!rc && *n != L.P->n) rc = fail(NEKBONE_ERR_ARG, "ax_: n differs from the bound mesh");
    if (!rc) rc = sync_geometry(gxyz, nullptr);
    if (!rc) rc = nekbone_ax(L.P, w, u);
    if (!rc && &cflops_) {
        const double nxyz = double(L.P->nx1) * L.P->nx1 * L.P->nx1;
        cflops_.flop_a += (19.0 * nxyz + 12.0 * L.P->nx1 * nxyz) * L.P->nelt;
    }
    report(rc, "ax_");
}

extern "C" void ax_e_(double *w, const double *u, const double *g, double *ur, double *us,
                      double *ut, double *wk)
{
    (void)ur, (void)us, (void)ut, (void)wk;
    LegacyState &L = S();
    int rc = bound("ax_e_");
    if (!rc && (!L.elem || L.elem->nx1 != L.P->nx1)) {
        if (L.elem) nekbone_problem_destroy(L.elem);
        L.elem = nullptr;
        const int nx = L.P->nx1;
        std::vector<int64_t> glo(size_t(nx) * nx * nx);
        for (size_t i = 0; i < glo.size(); ++i) glo[i] = int64_t(i) + 1;
        const int32_t one = 1;
        rc = nekbone_problem_create(&L.elem, 1, nx, 1, 1, 1, &one, L.P->dxm1, glo.data(),
                                    nullptr, nullptr);
    }
    if (!rc) rc = nekbone_set_geometry(L.elem, g);
    if (!rc) rc = nekbone_ax_local(L.elem, w, u);
    report(rc, "ax_e_");
}

static void grad_common(bool transpose, double *o0, double *o1, double *o2, double *out,
                        const double *in0, const double *in1, const double *in2, const int *n,
                        const double *D, const double *Dt)
{
    int rc = ensure_init();
    const int nx = *n + 1;
    const int64_t nxyz = int64_t(nx) * nx * nx;
    double *d = nullptr;
    if (!rc) rc = scratch(7 * nxyz + 2 * nx * nx, &d);
    if (rc) return report(rc, "local_grad3");
    double *dD = d + 7 * nxyz, *dDt = dD + nx * nx;
    const size_t eb = sizeof(double) * size_t(nxyz), db = sizeof(double) * size_t(nx * nx);
    bool ok = h2d(dD, D, db) == NEKBONE_OK &&
              h2d(dDt, Dt, db) == NEKBONE_OK;
    if (!transpose) {
        ok = ok && h2d(d, in0, eb) == NEKBONE_OK;
        hipLaunchKernelGGL(k_grad3, dim3(unsigned((nxyz + 255) / 256)), dim3(256), 0, stream(),
                           d + nxyz, d + 2 * nxyz, d + 3 * nxyz, d, dD, dDt, nx);
        (void)hipStreamSynchronize(stream());
        ok = ok && d2h(o0, d + nxyz, eb) == NEKBONE_OK &&
             d2h(o1, d + 2 * nxyz, eb) == NEKBONE_OK &&
             d2h(o2, d + 3 * nxyz, eb) == NEKBONE_OK;
    } else {
        ok = ok && h2d(d + nxyz, in0, eb) == NEKBONE_OK &&
             h2d(d + 2 * nxyz, in1, eb) == NEKBONE_OK &&
             h2d(d + 3 * nxyz, in2, eb) == NEKBONE_OK;
        hipLaunchKernelGGL(k_grad3_t, dim3(unsigned((nxyz + 255) / 256)), dim3(256), 0,
                           stream(), d, d + nxyz, d + 2 * nxyz, d + 3 * nxyz, dD, dDt, nx);
        (void)hipStreamSynchronize(stream());
        ok = ok && d2h(out, d, eb) == NEKBONE_OK;
    }
    if (!ok) report(fail(NEKBONE_ERR_HIP, "local_grad3 copy"), "local_grad3");
}

extern "C" void local_grad3_(double *ur, double *us, double *ut, const double *u, const int *n,
                             const double *D, const double *Dt)
{
    grad_common(false, ur, us, ut, nullptr, u, nullptr, nullptr, n, D, Dt);
}

extern "C" void local_grad3_t_(double *u, const double *ur, const double *us, const double *ut,
                               const int *N, const double *D, const double *Dt, double *w)
{
    (void)w;
    grad_common(true, nullptr, nullptr, nullptr, u, ur, us, ut, N, D, Dt);
}

extern "C" void dssum_(double *f)
{
    LegacyState &L = S();
    int rc = bound("dssum_");
    if (!rc) rc = nekbone_dssum(L.P, f);
    report(rc, "dssum_");
}

extern "C" void maskit_(double *w, double *pmask, const int *nx, const int *ny, const int *nz)
{
    (void)pmask, (void)nx, (void)ny, (void)nz;
    LegacyState &L = S();
    int rc = bound("maskit_");
    if (!rc) rc = nekbone_maskit(L.P, w);
    report(rc, "maskit_");
}

extern "C" double glsc3_(const double *a, const double *b, const double *mult, const int *n)
{
    const int64_t nn = *n;
    double *d = nullptr;
    int rc = ensure_init();
    if (!rc) rc = scratch(3 * nn + 2 * 1024 + 1, &d);
    if (rc) {
        report(rc, "glsc3_");
        return 0.0;
    }
    const size_t bytes = sizeof(double) * size_t(nn);
    double out = 0.0;
    nekbone_problem tmp;  // carries only n and the partials for the reduction
    tmp.n = nn;
    tmp.d_part = d + 3 * nn;
    bool ok = h2d(d, a, bytes) == NEKBONE_OK &&
              h2d(d + nn, b, bytes) == NEKBONE_OK &&
              h2d(d + 2 * nn, mult, bytes) == NEKBONE_OK;
    ok = ok && launch_glsc3(&tmp, d, d + nn, d + 2 * nn, d + 3 * nn + 2 * 1024) == NEKBONE_OK;
    (void)hipStreamSynchronize(stream());
    ok = ok && d2h(&out, d + 3 * nn + 2 * 1024, sizeof(double)) == NEKBONE_OK;
    tmp.d_part = nullptr;
    if (!ok) report(fail(NEKBONE_ERR_HIP, "glsc3_"), "glsc3_");
    return out;
}

extern "C" void gs_setup_(int *handle, const int64_t *glo_num, const int *n, const int *comm,
                          const int *np)
{
    (void)comm;
    LegacyState &L = S();
    if (np && *np != 1) {
        report(fail(NEKBONE_ERR_UNSUPPORTED, "gs_setup_: one rank per GPU (np must be 1)"),
               "gs_setup_");
        *handle = -1;
        return;
    }
    int rc = ensure_init();
    GsHandle h;
    h.n = *n;
    h.glo.assign(glo_num, glo_num + *n);
    std::vector<int32_t> off, idx;
    gs_groups(glo_num, *n, off, idx);
    h.map.ngroups = int64_t(off.size()) - 1;
    h.map.nshared = int64_t(idx.size());
    if (!rc) rc = dalloc(&h.map.d_off, int64_t(off.size()));
    if (!rc) rc = dalloc(&h.map.d_idx, int64_t(idx.size()));
    if (!rc) rc = dalloc(&h.d_u, int64_t(h.n));
    if (!rc) {
        rc = h2d(h.map.d_off, off.data(), sizeof(int32_t) * off.size());
        if (!rc) rc = h2d(h.map.d_idx, idx.data(), sizeof(int32_t) * idx.size());
    }
    if (rc) {
        report(fail(rc, "gs_setup_: device allocation"), "gs_setup_");
        *handle = -1;
        return;
    }
    *handle = L.next_handle++;
    L.gs[*handle] = std::move(h);
}

extern "C" void gs_op_(const int *handle, double *u, const int *dom, const int *op,
                       const int *transpose)
{
    (void)transpose;  // QQ^T is symmetric on one rank
    LegacyState &L = S();
    auto it = L.gs.find(*handle);
    if (it == L.gs.end()) return report(fail(NEKBONE_ERR_ARG, "gs_op_: invalid handle"), "gs_op_");
    // src/jl/gs.c:1178-1186 validates dom/op; this path implements
    // double (1) and add (1), the only combination the CG path uses.
    if (*dom != 1 || *op != 1)
        return report(fail(NEKBONE_ERR_UNSUPPORTED, "gs_op_: only dom=1 (double), op=1 (add)"),
                      "gs_op_");
    GsHandle &h = it->second;
    const size_t bytes = sizeof(double) * size_t(h.n);
    bool ok = h2d(h.d_u, u, bytes) == NEKBONE_OK;
    if (h.map.ngroups > 0)
        hipLaunchKernelGGL(k_gs_plain, dim3(unsigned((h.map.ngroups + 255) / 256)), dim3(256), 0,
                           stream(), h.d_u, h.map.d_off, h.map.d_idx, h.map.ngroups);
    (void)hipStreamSynchronize(stream());
    ok = ok && d2h(u, h.d_u, bytes) == NEKBONE_OK;
    if (!ok) report(fail(NEKBONE_ERR_HIP, "gs_op_ copy"), "gs_op_");
}

extern "C" void gs_free_(const int *handle)
{
    LegacyState &L = S();
    auto it = L.gs.find(*handle);
    if (it == L.gs.end()) return;
    GsHandle &h = it->second;
    dfree(h.map.d_off);
    dfree(h.map.d_idx);
    dfree(h.d_u);
    L.gs.erase(it);
}

// semhat (src/semhat.f:2-37): the 1-D spectral-element matrices of order
// n, column-major (0:n,0:n): a = Laplacian, b = GLL weights, c = b*d,
// d = derivative matrix, z = GLL points; w is the reference's work array
// (0:2n), unused here.  The reference host's proxy_setup (src/prox_setup.f:
// 2-25) copies d to dxm1, z to zgm1, b to wxm1.
extern "C" void semhat_(double *a, double *b, double *c, double *d, double *z, double *w,
                        const int *n)
{
    (void)w;
    if (!n || *n < 1 || *n + 1 > kMaxNx)
        return report(fail(NEKBONE_ERR_ARG, "semhat_: need 1 <= n <= 15"), "semhat_");
    semhat(a, b, c, d, z, *n);
}
