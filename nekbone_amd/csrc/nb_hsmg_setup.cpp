// HSMG preconditioner set-up on the host: h1mg_setup (src/hsmg.f:3-911)
// restated, then uploaded in the form the device kernels use.
//
// Every level quantity the reference computes is computed the same way
// (same numbering, same gather-scatter sums, same 1-D operators), with
// three changes of representation:
//  * The per-element FDM operators (S, S^T, eigenvalues) depend only on
//    the element's boundary type and neighbour widths in each direction,
//    so each distinct 1-D operator is stored once and elements keep three
//    indices; the diagonal d = 1/(lr+ls+lt) is formed in the kernel.
//  * The Schwarz and restriction weights are stored per point (1 where
//    the reference does not scale), and the Schwarz weights as their
//    square roots, the factor the reference applies.
//  * The ghost-layer exchange through the (n+2)^3 numbering is stored as
//    a gather table: for each ghost point of an element, the neighbour's
//    point whose value the exchange delivers there.
// The generalized eigenproblem (generalev, src/hsmg.f:483-523, LAPACK
// DSYGV in the reference) is solved by Cholesky reduction and cyclic
// Jacobi rotations; eigenvector signs may differ from LAPACK's, which the
// products S diag(d) S^T do not see.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "nb_internal.h"

namespace nb {
namespace {

constexpr int kM = kMaxNx + 2;  // largest extended size

struct HostGs {
    std::vector<int32_t> off, idx;
};

void gs_apply(const HostGs &G, double *u)
{
    for (size_t g = 0; g + 1 < G.off.size(); ++g) {
        const int32_t a = G.off[g], b = G.off[g + 1];
        double t = u[G.idx[size_t(a)]];
        for (int32_t q = a + 1; q < b; ++q) t += u[G.idx[size_t(q)]];
        for (int32_t q = a; q < b; ++q) u[G.idx[size_t(q)]] = t;
    }
}

HostGs gs_for(const nekbone_problem *P, int nx)
{
    std::vector<int64_t> glo;
    box_numbering(glo, nx, P->nelx, P->nely, P->nelz, P->lglel);
    HostGs G;
    gs_groups(glo.data(), int64_t(glo.size()), G.off, G.idx);
    return G;
}

// A x = lam B x, B symmetric positive definite; lam ascending, X^T B X = I
// (the definition DSYGV with itype=1 computes).  a (n x n, column-major)
// is overwritten by X.
void generalev(double *a, const double *b, double *lam, int n)
{
    double U[kM * kM] = {}, C[kM * kM], T[kM * kM], Y[kM * kM];
    for (int j = 0; j < n; ++j) {  // B = U^T U
        double s = b[j + n * j];
        for (int k = 0; k < j; ++k) s -= U[k + n * j] * U[k + n * j];
        U[j + n * j] = std::sqrt(s);
        for (int i = j + 1; i < n; ++i) {
            double t = b[j + n * i];
            for (int k = 0; k < j; ++k) t -= U[k + n * j] * U[k + n * i];
            U[j + n * i] = t / U[j + n * j];
        }
    }
    for (int i = 0; i < n; ++i)  // T = A U^-1
        for (int j = 0; j < n; ++j) {
            double t = a[i + n * j];
            for (int k = 0; k < j; ++k) t -= T[i + n * k] * U[k + n * j];
            T[i + n * j] = t / U[j + n * j];
        }
    for (int j = 0; j < n; ++j)  // C = U^-T T
        for (int i = 0; i < n; ++i) {
            double t = T[i + n * j];
            for (int k = 0; k < i; ++k) t -= U[k + n * i] * C[k + n * j];
            C[i + n * j] = t / U[i + n * i];
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i) {
            const double m = 0.5 * (C[i + n * j] + C[j + n * i]);
            C[i + n * j] = C[j + n * i] = m;
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) Y[i + n * j] = (i == j) ? 1.0 : 0.0;
    double fro = 0.0;
    for (int i = 0; i < n * n; ++i) fro += C[i] * C[i];
    for (int sweep = 0; sweep < 100; ++sweep) {  // cyclic Jacobi
        double off = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < j; ++i) off += C[i + n * j] * C[i + n * j];
        if (off <= 1e-30 * fro) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = C[p + n * q];
                if (apq == 0.0) continue;
                const double theta = (C[q + n * q] - C[p + n * p]) / (2.0 * apq);
                const double t =
                    (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double x = C[k + n * p], y = C[k + n * q];
                    C[k + n * p] = c * x - s * y;
                    C[k + n * q] = s * x + c * y;
                }
                for (int k = 0; k < n; ++k) {
                    const double x = C[p + n * k], y = C[q + n * k];
                    C[p + n * k] = c * x - s * y;
                    C[q + n * k] = s * x + c * y;
                }
                for (int k = 0; k < n; ++k) {
                    const double x = Y[k + n * p], y = Y[k + n * q];
                    Y[k + n * p] = c * x - s * y;
                    Y[k + n * q] = s * x + c * y;
                }
            }
    }
    int ord[kM];
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord, ord + n, [&](int x, int y) { return C[x * (n + 1)] < C[y * (n + 1)]; });
    for (int jj = 0; jj < n; ++jj) {
        const int j = ord[jj];
        lam[jj] = C[j * (n + 1)];
        for (int i = n - 1; i >= 0; --i) {  // x = U^-1 y
            double t = Y[i + n * j];
            for (int k = i + 1; k < n; ++k) t -= U[i + n * k] * a[k + n * jj];
            a[i + n * jj] = t / U[i + n * i];
        }
    }
}

// h1mg_setup_fast1d (src/hsmg.f:378-481): extended 1-D stiffness a and
// mass b of order n with neighbour widths ll, lr (element width lm), the
// generalized eigenvectors with Dirichlet rows zeroed (s) and their
// transpose (st), eigenvalues lam; nl = n+3.
void fast1d(double *s, double *st, double *lam, int lbc, int rbc, double ll, double lm, double lr,
            const double *ah, const double *bh, int n)
{
    const int m = n + 3;
    double b[kM * kM];
    auto A = [&](int i, int j) -> double & { return s[i + m * j]; };
    auto B = [&](int i, int j) -> double & { return b[i + m * j]; };
    auto AH = [&](int i, int j) { return ah[i + (n + 1) * j]; };
    const int i0 = (lbc == 1) ? 1 : 0, i1 = (rbc == 1) ? n - 1 : n;
    for (int i = 0; i < m * m; ++i) s[i] = b[i] = 0.0;
    double fac = 2.0 / lm;
    A(1, 1) = 1.0;
    A(n + 1, n + 1) = 1.0;
    for (int j = i0; j <= i1; ++j)
        for (int i = i0; i <= i1; ++i) A(i + 1, j + 1) = fac * AH(i, j);
    if (lbc == 0) {
        fac = 2.0 / ll;
        A(0, 0) = fac * AH(n - 1, n - 1);
        A(1, 0) = fac * AH(n, n - 1);
        A(0, 1) = fac * AH(n - 1, n);
        A(1, 1) = A(1, 1) + fac * AH(n, n);
    } else {
        A(0, 0) = 1.0;
    }
    if (rbc == 0) {
        fac = 2.0 / lr;
        A(n + 1, n + 1) = A(n + 1, n + 1) + fac * AH(0, 0);
        A(n + 2, n + 1) = fac * AH(1, 0);
        A(n + 1, n + 2) = fac * AH(0, 1);
        A(n + 2, n + 2) = fac * AH(1, 1);
    } else {
        A(n + 2, n + 2) = 1.0;
    }
    fac = 0.5 * lm;
    B(1, 1) = 1.0;
    B(n + 1, n + 1) = 1.0;
    for (int i = i0; i <= i1; ++i) B(i + 1, i + 1) = fac * bh[i];
    if (lbc == 0) {
        fac = 0.5 * ll;
        B(0, 0) = fac * bh[n - 1];
        B(1, 1) = B(1, 1) + fac * bh[n];
    } else {
        B(0, 0) = 1.0;
    }
    if (rbc == 0) {
        fac = 0.5 * lr;
        B(n + 1, n + 1) = B(n + 1, n + 1) + fac * bh[0];
        B(n + 2, n + 2) = fac * bh[1];
    } else {
        B(n + 2, n + 2) = 1.0;
    }
    generalev(s, b, lam, m);
    auto zero_row = [&](int r) {
        for (int j = 0; j < m; ++j) s[r + m * j] = 0.0;
    };
    if (lbc > 0) zero_row(0);
    if (lbc == 1) zero_row(1);
    if (rbc > 0) zero_row(m - 1);
    if (rbc == 1) zero_row(m - 2);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) st[i + m * j] = s[j + m * i];
}

// h1mg_extrude (src/hsmg.f:1246-1283) on extended arrays (nx^3 per element).
void extrude(double *a1, int l1, double f1, const double *a2, int l2, double f2, int nx, int nelt)
{
    const int64_t n3 = int64_t(nx) * nx * nx;
    auto ix = [nx](int i, int j, int k) { return i + nx * j + int64_t(nx) * nx * k; };
    for (int e = 0; e < nelt; ++e) {
        double *u = a1 + n3 * e;
        const double *v = a2 + n3 * e;
        for (int k = 1; k <= nx - 2; ++k)
            for (int j = 1; j <= nx - 2; ++j) {
                u[ix(l1, j, k)] = f1 * u[ix(l1, j, k)] + f2 * v[ix(l2, j, k)];
                u[ix(nx - 1 - l1, j, k)] = f1 * u[ix(nx - 1 - l1, j, k)] + f2 * v[ix(nx - 1 - l2, j, k)];
            }
        for (int k = 1; k <= nx - 2; ++k)
            for (int i = 1; i <= nx - 2; ++i) {
                u[ix(i, l1, k)] = f1 * u[ix(i, l1, k)] + f2 * v[ix(i, l2, k)];
                u[ix(i, nx - 1 - l1, k)] = f1 * u[ix(i, nx - 1 - l1, k)] + f2 * v[ix(i, nx - 1 - l2, k)];
            }
        for (int j = 1; j <= nx - 2; ++j)
            for (int i = 1; i <= nx - 2; ++i) {
                u[ix(i, j, l1)] = f1 * u[ix(i, j, l1)] + f2 * v[ix(i, j, l2)];
                u[ix(i, j, nx - 1 - l1)] = f1 * u[ix(i, j, nx - 1 - l1)] + f2 * v[ix(i, j, nx - 1 - l2)];
            }
    }
}

template <class T>
int upload(T **d, const std::vector<T> &h)
{
    if (int rc = dalloc(d, int64_t(h.size()))) return rc;
    return h2d(*d, h.data(), sizeof(T) * h.size());
}

// Ghost point of an extended element: face f (x-, x+, y-, y+, z-, z+) at
// tangential own index t = a + nh*b; ext coordinates.
void ghost_ext(int f, int t, int nh, int &I, int &J, int &K)
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

}  // namespace

void hsmg_free(nekbone_problem *P)
{
    if (!P->mg) return;
    drop_problem_graph(P);
    Hsmg *M = P->mg;
    for (int l = 1; l <= 3; ++l) {
        HsmgLevel &L = M->L[l];
        gs_map_free(L.gs);
        dfree(L.d_pos);
        dfree(L.d_shr);
        dfree(L.d_bnd);
        dfree(L.d_c);
        dfree(L.d_sw);
        dfree(L.d_gsrc);
        dfree(L.d_gin);
        dfree(L.d_gpart);
        dfree(L.d_ghost);
        dfree(L.d_ops);
        dfree(L.d_eop);
        dfree(L.d_eps);
        dfree(L.d_ops32);
        dfree(L.d_eps32);
        dfree(L.d_jh);
        dfree(L.d_jht);
        dfree(L.d_r);
        dfree(L.d_e);
        dfree(L.d_eown);
    }
    dfree(M->d_z);
    hsmg_coarse_free(M);
    delete M;
    P->mg = nullptr;
    P->hsmg_coarse = 0;
}

int hsmg_build(nekbone_problem *P)
{
    const int nx1 = P->nx1, nelt = P->nelt;
    if (nx1 < 4) return fail(NEKBONE_ERR_UNSUPPORTED, "hsmg: needs nx1 >= 4");
    hsmg_free(P);
    Hsmg *M = new Hsmg();
    P->mg = M;

    // h1mg_setup_mg_nx (src/hsmg.f:45-97).  For nx1 > 10 the reference
    // computes 2*(lx2/4)+1 from lx2, which its SIZE files leave undefined;
    // every lx2 >= 4 gives 3 after the min(3,.), the value taken here.
    M->lmax = (nx1 == 4) ? 2 : 3;
    static const int mgn2[10] = {1, 2, 2, 2, 2, 3, 3, 5, 5, 5};
    int mgnx2 = 3;
    if (nx1 <= 10) mgnx2 = mgn2[nx1 - 1];
    if (nx1 == 8) mgnx2 = 3;
    mgnx2 = std::min(3, mgnx2);
    int nxl[4] = {0, 1, mgnx2, mgnx2 + 1};
    nxl[M->lmax] = nx1 - 1;

    double ah[4][kM * kM], bh[4][kM], zh[4][kM];
    for (int l = 1; l <= M->lmax; ++l) {
        HsmgLevel &L = M->L[l];
        L.nx = nxl[l];
        L.nh = L.nx + 1;
        L.nl = L.nh + 2;
        L.n = int64_t(L.nh) * L.nh * L.nh * nelt;
        L.fine = (l == M->lmax);
        double c[kM * kM], d[kM * kM];
        semhat(ah[l], bh[l], c, d, zh[l], L.nx);  // h1mg_setup_semhat (:99-118)
    }
    // h1mg_setup_intp (:120-152): jh(nf,nc) Lagrange weights of the coarse
    // points at the fine ones
    for (int l = 1; l < M->lmax; ++l) {
        const int nf = M->L[l + 1].nh, nc = M->L[l].nh;
        std::vector<double> jh(size_t(nf) * nc), jht(size_t(nc) * nf);
        double w[2 * kM + 2];
        for (int i = 0; i < nf; ++i) {
            fd_weights_full(zh[l + 1][i], zh[l], nc - 1, 1, w);
            for (int j = 0; j < nc; ++j) jh[size_t(i + nf * j)] = w[j];
        }
        for (int i = 0; i < nf; ++i)
            for (int j = 0; j < nc; ++j) jht[size_t(j + nc * i)] = jh[size_t(i + nf * j)];
        if (int rc = upload(&M->L[l].d_jh, jh)) return rc;
        if (int rc = upload(&M->L[l].d_jht, jht)) return rc;
    }

    // Numbering and gather-scatter of every level (:154-169), masks
    // (:824-911), restriction weights = inverse multiplicity (:171-249).
    std::vector<HostGs> hgs(size_t(M->lmax) + 1);
    for (int l = 1; l <= M->lmax; ++l) {
        HsmgLevel &L = M->L[l];
        const int nh = L.nh;
        const int n3 = nh * nh * nh;
        hgs[size_t(l)] = gs_for(P, nh);
        std::vector<int32_t> pos(static_cast<size_t>(n3));
        for (int p = 0; p < n3; ++p) pos[size_t(p)] = L.fine ? P->perm[size_t(p)] : p;
        if (int rc = upload(&L.d_pos, pos)) return rc;
        if (!L.fine) {
            if (int rc = gs_map_upload(L.gs, hgs[size_t(l)].off, hgs[size_t(l)].idx, nullptr, n3))
                return rc;
            const HostGs &H = hgs[size_t(l)];
            std::vector<uint8_t> shr(size_t(L.n), 0);
            for (size_t g = 0; g + 1 < H.off.size(); ++g)
                if (H.off[g + 1] - H.off[g] >= 2)
                    for (int32_t q = H.off[g]; q < H.off[g + 1]; ++q) shr[size_t(H.idx[size_t(q)])] = 1;
            if (int rc = upload(&L.d_shr, shr)) return rc;
        }
        auto mem = [&](int64_t ref) -> int64_t {
            const int64_t e = ref / n3;
            return e * n3 + pos[size_t(ref - e * n3)];
        };
        std::vector<double> mult(size_t(L.n), 1.0), cinv(static_cast<size_t>(L.n));
        gs_apply(hgs[size_t(l)], mult.data());
        for (int64_t i = 0; i < L.n; ++i) cinv[size_t(mem(i))] = 1.0 / mult[size_t(i)];
        if (int rc = upload(&L.d_c, cinv)) return rc;
        std::vector<uint8_t> bnd(size_t(L.n), 0);
        for (int e = 0; e < nelt; ++e) {
            bool f[6];
            boundary_faces(P->lglel[size_t(e)], P->nelx, P->nely, P->nelz, f);
            for (int k = 0; k < nh; ++k)
                for (int j = 0; j < nh; ++j)
                    for (int i = 0; i < nh; ++i)
                        if (on_face(f, i, j, k, nh))
                            bnd[size_t(mem(int64_t(e) * n3 + i + nh * j + nh * nh * k))] = 1;
        }
        if (int rc = upload(&L.d_bnd, bnd)) return rc;
        if (!L.fine) {
            if (int rc = dalloc(&L.d_r, L.n)) return rc;
            if (int rc = dalloc(&L.d_e, L.n)) return rc;
        }
    }

    // Neighbour widths through the fine-level dssum (h1mg_setup_fast,
    // :300-331): width 2 of every element; a face without a neighbour
    // keeps 0.
    const double lm = 2.0;
    std::vector<double> lw(size_t(P->n), 0.0);
    {
        const int64_t f3 = int64_t(nx1) * nx1 * nx1;
        auto at = [&](int i, int j, int k, int e) -> double & {
            return lw[size_t(i + nx1 * j + int64_t(nx1) * nx1 * k + f3 * e)];
        };
        for (int e = 0; e < nelt; ++e) {
            at(0, 1, 1, e) = lm;
            at(nx1 - 1, 1, 1, e) = lm;
            at(1, 0, 1, e) = lm;
            at(1, nx1 - 1, 1, e) = lm;
            at(1, 1, 0, e) = lm;
            at(1, 1, nx1 - 1, e) = lm;
        }
        gs_apply(hgs[size_t(M->lmax)], lw.data());
        std::vector<double> w6(size_t(6) * nelt);
        for (int e = 0; e < nelt; ++e) {
            w6[size_t(6 * e + 0)] = at(0, 1, 1, e) - lm;
            w6[size_t(6 * e + 1)] = at(nx1 - 1, 1, 1, e) - lm;
            w6[size_t(6 * e + 2)] = at(1, 0, 1, e) - lm;
            w6[size_t(6 * e + 3)] = at(1, nx1 - 1, 1, e) - lm;
            w6[size_t(6 * e + 4)] = at(1, 1, 0, e) - lm;
            w6[size_t(6 * e + 5)] = at(1, 1, nx1 - 1, e) - lm;
        }
        lw.swap(w6);
    }

    for (int l = 2; l <= M->lmax; ++l) {
        HsmgLevel &L = M->L[l];
        const int nh = L.nh, nl = L.nl, n3 = nh * nh * nh, l3 = nl * nl * nl, g2 = 6 * nh * nh;
        std::vector<int32_t> pos(static_cast<size_t>(n3));
        for (int p = 0; p < n3; ++p) pos[size_t(p)] = L.fine ? P->perm[size_t(p)] : p;
        auto mem = [&](int e, int i, int j, int k) -> int32_t {
            return int32_t(int64_t(e) * n3 + pos[size_t(i + nh * j + nh * nh * k)]);
        };

        // FDM operators (:251-376), one per distinct boundary type and
        // neighbour widths
        std::map<std::tuple<int, int, double, double>, int> opid;
        std::vector<double> ops;
        std::vector<int16_t> eop(size_t(3) * nelt);
        std::vector<double> eps(static_cast<size_t>(nelt));
        std::vector<std::vector<double>> lams;
        for (int e = 0; e < nelt; ++e) {
            bool f[6];
            boundary_faces(P->lglel[size_t(e)], P->nelx, P->nely, P->nelz, f);  // get_bcs
            double mx[3];
            for (int d = 0; d < 3; ++d) {
                const int lbc = f[2 * d] ? 1 : 0, rbc = f[2 * d + 1] ? 1 : 0;
                const auto key = std::make_tuple(lbc, rbc, lw[size_t(6 * e + 2 * d)],
                                                 lw[size_t(6 * e + 2 * d + 1)]);
                auto it = opid.find(key);
                int id;
                if (it == opid.end()) {
                    id = int(opid.size());
                    opid[key] = id;
                    std::vector<double> S(size_t(nl) * nl), St(size_t(nl) * nl), lam(static_cast<size_t>(nl));
                    fast1d(S.data(), St.data(), lam.data(), lbc, rbc, std::get<2>(key), lm,
                           std::get<3>(key), ah[l], bh[l], L.nx);
                    ops.insert(ops.end(), S.begin(), S.end());
                    ops.insert(ops.end(), St.begin(), St.end());
                    ops.insert(ops.end(), lam.begin(), lam.end());
                    lams.push_back(lam);
                } else {
                    id = it->second;
                }
                eop[size_t(3 * e + d)] = int16_t(id);
                // vlmax(lam(2), nl-2): largest of eigenvalues 2..nl-1
                const std::vector<double> &lam = lams[size_t(id)];
                double m = lam[1];
                for (int i = 2; i <= nl - 2; ++i) m = std::max(m, lam[size_t(i)]);
                mx[d] = m;
            }
            eps[size_t(e)] = 1.e-5 * (mx[0] + mx[1] + mx[2]);
        }
        L.nops = int(opid.size());
        if (int rc = upload(&L.d_ops, ops)) return rc;
        if (int rc = upload(&L.d_eop, eop)) return rc;
        if (int rc = upload(&L.d_eps, eps)) return rc;
        if (int rc = upload(&L.d_ops32, std::vector<float>(ops.begin(), ops.end()))) return rc;
        if (int rc = upload(&L.d_eps32, std::vector<float>(eps.begin(), eps.end()))) return rc;

        // Extended numbering (nh+2 points, :154-169): each face ghost
        // point is shared with exactly one ghost of the neighbour.
        HostGs xg = gs_for(P, nl);
        std::vector<int32_t> partner(size_t(l3) * nelt, -1);
        for (size_t g = 0; g + 1 < xg.off.size(); ++g)
            if (xg.off[g + 1] - xg.off[g] == 2) {
                const int32_t a = xg.idx[size_t(xg.off[g])], b = xg.idx[size_t(xg.off[g] + 1)];
                partner[size_t(a)] = b;
                partner[size_t(b)] = a;
            }
        std::vector<int32_t> gsrc(size_t(g2) * nelt, -1), gin(size_t(g2) * nelt),
            gpart(size_t(g2) * nelt, -1);
        for (int e = 0; e < nelt; ++e)
            for (int f = 0; f < 6; ++f)
                for (int t = 0; t < nh * nh; ++t) {
                    int I, J, K;
                    ghost_ext(f, t, nh, I, J, K);
                    const size_t slot = size_t(int64_t(e) * g2 + f * nh * nh + t);
                    // own point one inside this face: ext layer 2 / nl-3
                    const int di = (f == 0) ? 2 : (f == 1) ? -2 : 0;
                    const int dj = (f == 2) ? 2 : (f == 3) ? -2 : 0;
                    const int dk = (f == 4) ? 2 : (f == 5) ? -2 : 0;
                    gin[slot] = mem(e, I + di - 1, J + dj - 1, K + dk - 1);
                    const int32_t q = partner[size_t(int64_t(e) * l3 + I + nl * J + nl * nl * K)];
                    if (q < 0) continue;
                    const int e2 = q / l3, r = q - e2 * l3;
                    const int I2 = r % nl, J2 = (r / nl) % nl, K2 = r / (nl * nl);
                    int f2, a2, b2;
                    if (I2 == 0 || I2 == nl - 1) {
                        f2 = I2 ? 1 : 0;
                        a2 = J2 - 1;
                        b2 = K2 - 1;
                    } else if (J2 == 0 || J2 == nl - 1) {
                        f2 = J2 ? 3 : 2;
                        a2 = I2 - 1;
                        b2 = K2 - 1;
                    } else {
                        f2 = K2 ? 5 : 4;
                        a2 = I2 - 1;
                        b2 = J2 - 1;
                    }
                    // the neighbour contributes its point two in from its ghost
                    const int ei = (f2 == 0) ? 2 : (f2 == 1) ? -2 : 0;
                    const int ej = (f2 == 2) ? 2 : (f2 == 3) ? -2 : 0;
                    const int ek = (f2 == 4) ? 2 : (f2 == 5) ? -2 : 0;
                    gsrc[slot] = mem(e2, I2 + ei - 1, J2 + ej - 1, K2 + ek - 1);
                    gpart[slot] = int32_t(int64_t(e2) * g2 + f2 * nh * nh + a2 + nh * b2);
                }
        if (int rc = upload(&L.d_gsrc, gsrc)) return rc;
        if (int rc = upload(&L.d_gin, gin)) return rc;
        if (int rc = upload(&L.d_gpart, gpart)) return rc;
        if (int rc = dalloc(&L.d_ghost, int64_t(g2) * nelt)) return rc;
        if (int rc = dalloc(&L.d_eown, L.n)) return rc;

        // Schwarz weights (h1mg_setup_schwarz_wt3d_1, :606-652) and the
        // factors h1mg_schwarz_wt3d2 (:1144-1178) applies
        const int64_t ns = int64_t(l3) * nelt;
        std::vector<double> w0(size_t(ns), 0.0), wi(size_t(ns), 1.0);
        extrude(w0.data(), 0, 0.0, wi.data(), 0, 1.0, nl, nelt);
        gs_apply(xg, wi.data());
        extrude(wi.data(), 0, 1.0, w0.data(), 0, -1.0, nl, nelt);
        extrude(wi.data(), 2, 1.0, wi.data(), 0, 1.0, nl, nelt);
        std::vector<double> wr(static_cast<size_t>(L.n));
        for (int e = 0; e < nelt; ++e)
            for (int k = 0; k < nh; ++k)
                for (int j = 0; j < nh; ++j)
                    for (int i = 0; i < nh; ++i)
                        wr[size_t(int64_t(e) * n3 + i + nh * j + nh * nh * k)] =
                            wi[size_t(int64_t(e) * l3 + (i + 1) + nl * (j + 1) + nl * nl * (k + 1))];
        gs_apply(hgs[size_t(l)], wr.data());
        std::vector<double> sw(size_t(L.n), 1.0);
        auto W = [&](int e, int i, int j, int k) {
            return wr[size_t(int64_t(e) * n3 + i + nh * j + nh * nh * k)];
        };
        auto scale = [&](int e, int i, int j, int k, double wt) {
            double &v = sw[size_t(mem(e, i, j, k))];
            v = v * std::sqrt(wt);
        };
        const int lay[4] = {0, 1, nh - 2, nh - 1};
        for (int e = 0; e < nelt; ++e) {
            for (int k = 0; k < nh; ++k)
                for (int j = 0; j < nh; ++j)
                    for (int q = 0; q < 4; ++q) scale(e, lay[q], j, k, 1.0 / W(e, lay[q], j, k));
            for (int k = 0; k < nh; ++k)
                for (int i = 2; i <= nh - 3; ++i)
                    for (int q = 0; q < 4; ++q) scale(e, i, lay[q], k, 1.0 / W(e, i, lay[q], k));
            for (int j = 2; j <= nh - 3; ++j)
                for (int i = 2; i <= nh - 3; ++i)
                    for (int q = 0; q < 4; ++q) scale(e, i, j, lay[q], 1.0 / W(e, i, j, lay[q]));
        }
        if (int rc = upload(&L.d_sw, sw)) return rc;
    }
    if (int rc = dalloc(&M->d_z, P->n)) return rc;
    if (int rc = dzero(M->d_z, sizeof(double) * size_t(P->n))) return rc;
    return NEKBONE_OK;
}

}  // namespace nb
