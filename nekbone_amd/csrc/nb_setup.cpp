// Host-side set-up of a box problem: GLL rule, 1-D SEM operators, global
// numbering, gather-scatter groups, Dirichlet lists.  These run once per
// mesh; everything per-iteration runs on the device.
//
// The GLL rule and the derivative matrix are produced with exactly the
// reference's floating-point operations (same formulas, same order), so D
// is bit-identical to the reference's dxm1 and the element operator built
// on it can be compared bit for bit.
#include <algorithm>
#include <cmath>

#include "nb_internal.h"

namespace nb {
namespace {

// Gamma function at the half-integer / integer points the GLL rule needs,
// src/speclib.f:371-393.
double gamma_table(double x)
{
    const double pi = 4.0 * std::atan(1.0);
    double g = 1.0;
    if (x == -0.5) g = -2.0 * std::sqrt(pi);
    if (x == 0.5) g = std::sqrt(pi);
    if (x == 1.0) g = 1.0;
    if (x == 2.0) g = 1.0;
    if (x == 1.5) g = std::sqrt(pi) / 2.;
    if (x == 2.5) g = 1.5 * std::sqrt(pi) / 2.;
    if (x == 3.5) g = 0.5 * (2.5 * (1.5 * std::sqrt(pi)));
    if (x == 3.) g = 2.;
    if (x == 4.) g = 6.;
    if (x == 5.) g = 24.;
    if (x == 6.) g = 120.;
    return g;
}

struct JacobiEval {
    double p, pd, pm1, pdm1, pm2, pdm2;
};

// Jacobi polynomial P_n^(a,b)(x) and derivative by the three-term
// recurrence, src/speclib.f:479-518.
JacobiEval jacobi_poly(int n, double alp, double bet, double x)
{
    JacobiEval r{1.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (n == 0) return r;
    const double apb = alp + bet;
    double polyl = r.p, pderl = r.pd, psave = 0.0, pdsave = 0.0;
    r.p = (alp - bet + (apb + 2.) * x) / 2.;
    r.pd = (apb + 2.) / 2.;
    if (n == 1) return r;
    for (int k = 2; k <= n; ++k) {
        const double dk = k;
        const double a1 = 2. * dk * (dk + apb) * (2. * dk + apb - 2.);
        const double a2 = (2. * dk + apb - 1.) * (alp * alp - bet * bet);
        const double b3 = (2. * dk + apb - 2.);
        const double a3 = b3 * (b3 + 1.) * (b3 + 2.);
        const double a4 = 2. * (dk + alp - 1.) * (dk + bet - 1.) * (2. * dk + apb);
        const double pn = ((a2 + a3 * x) * r.p - a4 * polyl) / a1;
        const double pdn = ((a2 + a3 * x) * r.pd - a4 * pderl + a3 * r.p) / a1;
        psave = polyl;
        pdsave = pderl;
        polyl = r.p;
        r.p = pn;
        pderl = r.pd;
        r.pd = pdn;
    }
    r.pm1 = polyl;
    r.pdm1 = pderl;
    r.pm2 = psave;
    r.pdm2 = pdsave;
    return r;
}

// Zeros of P_np^(a,b): Newton with deflation from Chebyshev-like starts,
// then a selection sort, src/speclib.f:421-477.
void jacobi_zeros(double *xj, int np, double alpha, double beta)
{
    const int n = np - 1;
    const double dth = 4. * std::atan(1.0) / (2. * double(n) + 2.);
    double xlast = 0.0;
    for (int j = 1; j <= np; ++j) {
        double x = std::cos((2. * (double(j) - 1.) + 1.) * dth);
        if (j > 1) x = (x + xlast) / 2.;
        for (int it = 0; it < 10; ++it) {
            const JacobiEval e = jacobi_poly(np, alpha, beta, x);
            double recsum = 0.;
            for (int i = 1; i <= j - 1; ++i) recsum = recsum + 1. / (x - xj[np - i]);
            const double delx = -e.p / (e.pd - recsum * e.p);
            x = x + delx;
            if (std::fabs(delx) < 1.0e-12) break;
        }
        xj[np - j] = x;
        xlast = x;
    }
    for (int i = 0; i < np; ++i) {
        double xmin = 2.;
        int jmin = i;
        for (int j = i; j < np; ++j)
            if (xj[j] < xmin) {
                xmin = xj[j];
                jmin = j;
            }
        if (jmin != i) std::swap(xj[i], xj[jmin]);
    }
}

// Normalisation integral, src/speclib.f:395-419.
double jacobi_norm(int n, double alpha, double beta)
{
    const double dn = n, cnst = alpha + beta + 1.;
    double prod;
    if (n <= 1) {
        prod = gamma_table(dn + alpha) * gamma_table(dn + beta);
        prod = prod / (gamma_table(dn) * gamma_table(dn + alpha + beta));
        return prod * std::pow(2., cnst) / (2. * dn + cnst);
    }
    prod = gamma_table(alpha + 1.) * gamma_table(beta + 1.);
    prod = prod / (2. * (1. + cnst) * gamma_table(cnst + 1.));
    prod = prod * (1. + alpha) * (2. + alpha);
    prod = prod * (1. + beta) * (2. + beta);
    for (int i = 3; i <= n; ++i) {
        const double di = i;
        const double frac = (di + alpha) * (di + beta) / (di * (di + alpha + beta));
        prod = prod * frac;
    }
    return prod * std::pow(2., cnst) / (2. * dn + cnst);
}

// Gauss-Jacobi rule, src/speclib.f:155-205.
void gauss_jacobi(double *z, double *w, int np, double alpha, double beta)
{
    const double apb = alpha + beta;
    if (np == 1) {
        z[0] = (beta - alpha) / (apb + 2.);
        w[0] = gamma_table(alpha + 1.) * gamma_table(beta + 1.) / gamma_table(apb + 2.) *
               std::pow(2., apb + 1.);
        return;
    }
    jacobi_zeros(z, np, alpha, beta);
    const int n = np - 1;
    const double dnp1 = n + 1, dnp2 = n + 2;
    const double fac1 = dnp1 + alpha + beta + 1.;
    const double fac2 = fac1 + dnp1;
    const double fac3 = fac2 + 1.;
    const double fnorm = jacobi_norm(n + 1, alpha, beta);
    const double rcoef = (fnorm * fac2 * fac3) / (2. * fac1 * dnp2);
    for (int i = 0; i < np; ++i) {
        const JacobiEval e = jacobi_poly(n + 2, alpha, beta, z[i]);
        w[i] = -rcoef / (e.p * e.pdm1);
    }
}

// End-point weights of the Gauss-Lobatto-Jacobi rule,
// src/speclib.f:283-369 (left = ENDW1, right = ENDW2).
double lobatto_end_weight(int n, double alpha, double beta, bool left)
{
    const double apb = alpha + beta;
    if (n == 0) return 0.;
    // ENDW1 and ENDW2 differ by which of (alpha, beta) gets the +1 shift
    // in the first terms and by the sign in f2.
    const double ga = left ? alpha + 2. : alpha + 1.;
    const double gb = left ? beta + 1. : beta + 2.;
    double f1 = gamma_table(ga) * gamma_table(gb) / gamma_table(apb + 3.);
    f1 = f1 * (apb + 2.) * std::pow(2., apb + 2.) / 2.;
    if (n == 1) return f1;
    double fint1 = gamma_table(ga) * gamma_table(gb) / gamma_table(apb + 3.);
    fint1 = fint1 * std::pow(2., apb + 2.);
    double fint2 = gamma_table(alpha + 2.) * gamma_table(beta + 2.) / gamma_table(apb + 4.);
    fint2 = fint2 * std::pow(2., apb + 3.);
    double f2 = left ? (-2. * (beta + 2.) * fint1 + (apb + 4.) * fint2) * (apb + 3.) / 4.
                     : (2. * (alpha + 2.) * fint1 - (apb + 4.) * fint2) * (apb + 3.) / 4.;
    if (n == 2) return f2;
    double f3 = 0.;
    for (int i = 3; i <= n; ++i) {
        const double di = i - 1;
        const double abn = alpha + beta + di, abnn = abn + di;
        const double a1 = -(2. * (di + alpha) * (di + beta)) / (abn * abnn * (abnn + 1.));
        const double a2 = (2. * (alpha - beta)) / (abnn * (abnn + 2.));
        const double a3 = (2. * (abn + 1.)) / ((abnn + 2.) * (abnn + 1.));
        f3 = -(a2 * f2 + a1 * f1) / a3;
        f1 = f2;
        f2 = f3;
    }
    return f3;
}

int mod1(int i, int n)
{
    if (i == 0 || n == 0) return 0;
    return (i + n - 1) % n + 1;
}

}  // namespace

// Gauss-Lobatto-Legendre rule: interior nodes are the Gauss-Jacobi(1,1)
// nodes, src/speclib.f:107-122 -> 238-281.
void gll_points_weights(double *z, double *w, int np)
{
    const int n = np - 1;
    if (n - 1 > 0) gauss_jacobi(z + 1, w + 1, n - 1, 1.0, 1.0);
    z[0] = -1.;
    z[np - 1] = 1.;
    for (int i = 1; i < np - 1; ++i) w[i] = w[i] / (1. - z[i] * z[i]);
    JacobiEval e = jacobi_poly(n, 0., 0., z[0]);
    w[0] = lobatto_end_weight(n, 0., 0., true) / (2. * e.pd);
    e = jacobi_poly(n, 0., 0., z[np - 1]);
    w[np - 1] = lobatto_end_weight(n, 0., 0., false) / (2. * e.pd);
}

// Fornberg finite-difference weights on all points, src/semhat.f:39-93.
void fd_weights_full(double xx, const double *x, int n, int m, double *c)
{
    auto C = [&](int j, int k) -> double & { return c[j + (n + 1) * k]; };
    double c1 = 1., c4 = x[0] - xx;
    for (int k = 0; k <= m; ++k)
        for (int j = 0; j <= n; ++j) C(j, k) = 0.;
    C(0, 0) = 1.;
    for (int i = 1; i <= n; ++i) {
        const int mn = std::min(i, m);
        double c2 = 1.;
        const double c5 = c4;
        c4 = x[i] - xx;
        for (int j = 0; j < i; ++j) {
            const double c3 = x[i] - x[j];
            c2 = c2 * c3;
            for (int k = mn; k >= 1; --k)
                C(i, k) = c1 * (double(k) * C(i - 1, k - 1) - c5 * C(i - 1, k)) / c2;
            C(i, 0) = -c1 * c5 * C(i - 1, 0) / c2;
            for (int k = mn; k >= 1; --k) C(j, k) = (c4 * C(j, k) - double(k) * C(j, k - 1)) / c3;
            C(j, 0) = c4 * C(j, 0) / c3;
        }
        c1 = c2;
    }
}

// 1-D SEM matrices: a = D^T B D, b = weights, c = B D, d = D, z = nodes,
// src/semhat.f:2-37.  Column-major (n+1)x(n+1).
void semhat(double *a, double *b, double *c, double *d, double *z, int n)
{
    const int np = n + 1;
    double wk[2 * kMaxNx + 4];
    gll_points_weights(z, b, np);
    for (int i = 0; i <= n; ++i) {
        fd_weights_full(z[i], z, n, 1, wk);
        for (int j = 0; j <= n; ++j) d[i + np * j] = wk[j + np];
    }
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
            double s = 0.;
            for (int k = 0; k <= n; ++k) s = s + d[k + np * i] * b[k] * d[k + np * j];
            a[i + np * j] = s;
            c[i + np * j] = b[i] * d[i + np * j];
        }
}

// Global node numbers of a box, src/prox_dssum.f:42-86.
void box_numbering(std::vector<int64_t> &glo, int nx, int nelx, int nely, int nelz,
                   const std::vector<int32_t> &lglel)
{
    (void)nelz;
    const int nn = nx - 1;
    const int64_t nxyz = int64_t(nx) * nx * nx;
    const int64_t sx = int64_t(nn) * nelx + 1, sxy = sx * (int64_t(nn) * nely + 1);
    glo.resize(nxyz * lglel.size());
    for (size_t e = 0; e < lglel.size(); ++e) {
        const int eg = lglel[e], nelxy = nelx * nely;
        const int ez = 1 + (eg - 1) / nelxy;
        const int ey = 1 + (mod1(eg, nelxy) - 1) / nelx;
        const int ex = mod1(eg, nelx);
        int64_t *o = glo.data() + nxyz * e;
        for (int k = 0; k <= nn; ++k)
            for (int j = 0; j <= nn; ++j)
                for (int i = 0; i <= nn; ++i) {
                    const int64_t kg = int64_t(nn) * (ez - 1) + k;
                    const int64_t jg = int64_t(nn) * (ey - 1) + j;
                    const int64_t ig = int64_t(nn) * (ex - 1) + i;
                    o[i + nx * j + nx * nx * k] = 1 + ig + jg * sx + kg * sxy;
                }
    }
}

// Shared-node groups in the jl local order.  A counting sort by id,
// filled in ascending local index, reproduces the stable sort of
// src/jl/gs.c:105-112; emitting a group when its first member is met in
// an ascending sweep orders groups by primary (gs.c:112, 330-357).
void gs_groups(const int64_t *glo, int64_t n, std::vector<int32_t> &off,
               std::vector<int32_t> &idx)
{
    int64_t maxid = 0;
    for (int64_t l = 0; l < n; ++l) maxid = std::max(maxid, glo[l]);
    std::vector<int64_t> start(size_t(maxid) + 2, 0);
    for (int64_t l = 0; l < n; ++l)
        if (glo[l] > 0) start[size_t(glo[l]) + 1]++;
    for (int64_t i = 1; i <= maxid + 1; ++i) start[size_t(i)] += start[size_t(i - 1)];
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    std::vector<int32_t> byid(size_t(std::max<int64_t>(n, 1)));
    for (int64_t l = 0; l < n; ++l)
        if (glo[l] > 0) byid[size_t(fill[size_t(glo[l])]++)] = int32_t(l);
    off.clear();
    idx.clear();
    off.push_back(0);
    for (int64_t l = 0; l < n; ++l) {
        const int64_t id = glo[l];
        if (id <= 0) continue;
        const int64_t a = start[size_t(id)], b = start[size_t(id) + 1];
        if (b - a >= 2 && byid[size_t(a)] == l) {
            for (int64_t q = a; q < b; ++q) idx.push_back(byid[size_t(q)]);
            off.push_back(int32_t(idx.size()));
        }
    }
}

// Device point order inside an element: the points on the element's
// surface first (in reference order), then the interior points (in
// reference order).  perm[p] = device position of reference point
// p = i + nx*j + nx*nx*k.  All shared points lie on element surfaces, so
// the gather-scatter touches only the dense leading block of each element.
void boundary_first_perm(int nx, std::vector<int32_t> &perm)
{
    const int n3 = nx * nx * nx;
    perm.assign(size_t(n3), -1);
    for (int p = 0; p < n3; ++p) {
        const int i = p % nx, j = (p / nx) % nx, k = p / (nx * nx);
        perm[size_t(p)] = surface_order_pos(nx, i, j, k);
    }
}

bool surface_order_consistent(int nx)
{
    const int n3 = nx * nx * nx;
    std::vector<int> seen(size_t(n3), 0);
    for (int p = 0; p < n3; ++p) {
        const int d = surface_order_pos(nx, p % nx, (p / nx) % nx, p / (nx * nx));
        if (d < 0 || d >= n3 || seen[size_t(d)]++ || surface_order_ref(nx, d) != p) return false;
    }
    return true;
}

// Faces of element eg (1-based global number) on the domain boundary, as
// get_face decides, src/cg.f:302-368: {x0 (Face4), x1 (Face2), y0 (Face1),
// y1 (Face3), z0 (Face5), z1 (Face6)}.
void boundary_faces(int eg, int nelx, int nely, int nelz, bool f[6])
{
    const int nelxy = nelx * nely;
    const int ir = 1 + (eg - 1) / nelxy;
    const int iq = 1 + (mod1(eg, nelxy) - 1) / nelx;
    const int ip = mod1(eg, nelx);
    f[0] = (ip % nelx == 1) || nelx == 1;
    f[1] = (ip % nelx == 0);
    f[2] = (iq % nely == 1) || nely == 1;
    f[3] = (iq % nely == 0);
    f[4] = (ir % nelz == 1) || nelz == 1;
    f[5] = (ir % nelz == 0);
}

bool on_face(const bool f[6], int i, int j, int k, int nx)
{
    return (f[0] && i == 0) || (f[1] && i == nx - 1) || (f[2] && j == 0) ||
           (f[3] && j == nx - 1) || (f[4] && k == 0) || (f[5] && k == nx - 1);
}

void mask_lists(int nx, int nelt, int nelx, int nely, int nelz,
                const std::vector<int32_t> &lglel, std::vector<int32_t> &full,
                std::vector<int32_t> &ref)
{
    const int64_t nxyz = int64_t(nx) * nx * nx;
    full.clear();
    ref.clear();
    for (int e = 0; e < nelt; ++e) {
        bool f[6];
        boundary_faces(lglel[size_t(e)], nelx, nely, nelz, f);
        for (int k = 0; k < nx; ++k)
            for (int j = 0; j < nx; ++j)
                for (int i = 0; i < nx; ++i)
                    if (on_face(f, i, j, k, nx)) {
                        const int32_t l = int32_t(i + nx * j + nx * nx * k + nxyz * e);
                        full.push_back(l);
                        if (e == 0) ref.push_back(l);
                    }
    }
}

}  // namespace nb
