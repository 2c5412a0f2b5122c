/* Oracle: CPU restatement of Nekbone's CG Poisson hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see nekbone_oracle.h).  Built with
 * -ffp-contract=off and no fast-math so that every expression is evaluated
 * with the same roundings, in the same order, as the flang-compiled
 * reference (x86-64 baseline, no FMA).  Element and point loops are
 * parallelised with OpenMP where each iteration is independent, which does
 * not change any result; the glsc3 sum stays strictly sequential like the
 * reference's.
 */
#include "nekbone_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define MAXNX 32

struct orc_problem {
    int nelx, nely, nelz, nelt, nx1;
    int64_t n;
    double z[MAXNX], w[MAXNX];
    double d[MAXNX * MAXNX];  /* dxm1(i,j) = d[i + nx*j] */
    double dt[MAXNX * MAXNX]; /* dxtm1 = transpose */
    double *g;                /* g(6,n) */
    double *c;                /* 1/multiplicity */
    double *dinv;             /* Jacobi */
    int64_t *glo;
    int64_t ngroups, nshared;
    int64_t *goff;
    int32_t *gidx;
    int64_t nfull, nref;
    int32_t *mfull, *mref;
    int mask_mode;
    int mask_built; /* 0: cmask(-1)>=0 (next maskit builds the list) */
    int exact_dots; /* 1: orc_cg's dot products correctly rounded (orc_glsc3_exact) */
    double *wk;     /* scratch n */
    struct orc_hsmg *mg; /* HSMG preconditioner, NULL until orc_hsmg_setup */
};
static void hsmg_free(struct orc_hsmg *M);

/* ------------------------------------------------------------------ */
/* GLL points and weights: src/speclib.f:107-518                        */
/* ------------------------------------------------------------------ */

/* GAMMAF, src/speclib.f:371-393 */
static double gammaf(double x)
{
    const double half = 0.5, one = 1.0, two = 2.0, four = 4.0;
    double pi = four * atan(one);
    double g = one;
    if (x == -half) g = -two * sqrt(pi);
    if (x == half) g = sqrt(pi);
    if (x == one) g = one;
    if (x == two) g = one;
    if (x == 1.5) g = sqrt(pi) / 2.;
    if (x == 2.5) g = 1.5 * sqrt(pi) / 2.;
    if (x == 3.5) g = 0.5 * (2.5 * (1.5 * sqrt(pi)));
    if (x == 3.) g = 2.;
    if (x == 4.) g = 6.;
    if (x == 5.) g = 24.;
    if (x == 6.) g = 120.;
    return g;
}

/* PNORMJ, src/speclib.f:395-419 */
static double pnormj(int n, double alpha, double beta)
{
    const double one = 1., two = 2.;
    double dn = (double)n;
    double cnst = alpha + beta + one;
    double prod;
    if (n <= 1) {
        prod = gammaf(dn + alpha) * gammaf(dn + beta);
        prod = prod / (gammaf(dn) * gammaf(dn + alpha + beta));
        return prod * pow(two, cnst) / (two * dn + cnst);
    }
    prod = gammaf(alpha + one) * gammaf(beta + one);
    prod = prod / (two * (one + cnst) * gammaf(cnst + one));
    prod = prod * (one + alpha) * (two + alpha);
    prod = prod * (one + beta) * (two + beta);
    for (int i = 3; i <= n; ++i) {
        double dindx = (double)i;
        double frac = (dindx + alpha) * (dindx + beta) / (dindx * (dindx + alpha + beta));
        prod = prod * frac;
    }
    return prod * pow(two, cnst) / (two * dn + cnst);
}

/* JACOBF, src/speclib.f:479-518 */
static void jacobf(double *poly, double *pder, double *polym1, double *pderm1,
                   double *polym2, double *pderm2, int n, double alp,
                   double bet, double x)
{
    double apb = alp + bet;
    double p = 1., pd = 0.;
    double polyl, pderl, psave = 0., pdsave = 0.;
    *poly = p;
    *pder = pd;
    if (n == 0) return;
    polyl = p;
    pderl = pd;
    p = (alp - bet + (apb + 2.) * x) / 2.;
    pd = (apb + 2.) / 2.;
    *poly = p;
    *pder = pd;
    if (n == 1) return;
    for (int k = 2; k <= n; ++k) {
        double dk = (double)k;
        double a1 = 2. * dk * (dk + apb) * (2. * dk + apb - 2.);
        double a2 = (2. * dk + apb - 1.) * (alp * alp - bet * bet);
        double b3 = (2. * dk + apb - 2.);
        double a3 = b3 * (b3 + 1.) * (b3 + 2.);
        double a4 = 2. * (dk + alp - 1.) * (dk + bet - 1.) * (2. * dk + apb);
        double polyn = ((a2 + a3 * x) * p - a4 * polyl) / a1;
        double pdern = ((a2 + a3 * x) * pd - a4 * pderl + a3 * p) / a1;
        psave = polyl;
        pdsave = pderl;
        polyl = p;
        p = polyn;
        pderl = pd;
        pd = pdern;
    }
    *poly = p;
    *pder = pd;
    *polym1 = polyl;
    *pderm1 = pderl;
    *polym2 = psave;
    *pderm2 = pdsave;
}

/* JACG, src/speclib.f:421-477 (xjac is 1-based in the reference) */
static void jacg(double *xjac, int np, double alpha, double beta)
{
    const int kstop = 10;
    const double eps = 1.0e-12;
    int n = np - 1;
    double one = 1.;
    double dth = 4. * atan(one) / (2. * ((double)n) + 2.);
    double xlast = 0.;
    double p, pd, pm1, pdm1, pm2, pdm2;
    for (int j = 1; j <= np; ++j) {
        double x;
        if (j == 1) {
            x = cos((2. * (((double)j) - 1.) + 1.) * dth);
        } else {
            double x1 = cos((2. * (((double)j) - 1.) + 1.) * dth);
            double x2 = xlast;
            x = (x1 + x2) / 2.;
        }
        for (int k = 1; k <= kstop; ++k) {
            jacobf(&p, &pd, &pm1, &pdm1, &pm2, &pdm2, np, alpha, beta, x);
            double recsum = 0.;
            int jm = j - 1;
            for (int i = 1; i <= jm; ++i)
                recsum = recsum + 1. / (x - xjac[(np - i + 1) - 1]);
            double delx = -p / (pd - recsum * p);
            x = x + delx;
            if (fabs(delx) < eps) break;
        }
        xjac[(np - j + 1) - 1] = x;
        xlast = x;
    }
    for (int i = 1; i <= np; ++i) {
        double xmin = 2.;
        int jmin = i;
        for (int j = i; j <= np; ++j) {
            if (xjac[j - 1] < xmin) {
                xmin = xjac[j - 1];
                jmin = j;
            }
        }
        if (jmin != i) {
            double swap = xjac[i - 1];
            xjac[i - 1] = xjac[jmin - 1];
            xjac[jmin - 1] = swap;
        }
    }
}

/* ZWGJD, src/speclib.f:155-205 */
static void zwgjd(double *z, double *w, int np, double alpha, double beta)
{
    int n = np - 1;
    const double one = 1., two = 2.;
    double apb = alpha + beta;
    (void)n;
    if (np == 1) {
        z[0] = (beta - alpha) / (apb + two);
        w[0] = gammaf(alpha + one) * gammaf(beta + one) / gammaf(apb + two) *
               pow(two, apb + one);
        return;
    }
    jacg(z, np, alpha, beta);
    int np1 = n + 1, np2 = n + 2;
    double dnp1 = (double)np1, dnp2 = (double)np2;
    double fac1 = dnp1 + alpha + beta + one;
    double fac2 = fac1 + dnp1;
    double fac3 = fac2 + one;
    double fnorm = pnormj(np1, alpha, beta);
    double rcoef = (fnorm * fac2 * fac3) / (two * fac1 * dnp2);
    for (int i = 0; i < np; ++i) {
        double p, pd, pm1, pdm1, pm2, pdm2;
        jacobf(&p, &pd, &pm1, &pdm1, &pm2, &pdm2, np2, alpha, beta, z[i]);
        w[i] = -rcoef / (p * pdm1);
    }
}

/* ENDW1, src/speclib.f:283-325 */
static double endw1(int n, double alpha, double beta)
{
    const double zero = 0., one = 1., two = 2., three = 3., four = 4.;
    double apb = alpha + beta;
    if (n == 0) return zero;
    double f1 = gammaf(alpha + two) * gammaf(beta + one) / gammaf(apb + three);
    f1 = f1 * (apb + two) * pow(two, apb + two) / two;
    if (n == 1) return f1;
    double fint1 = gammaf(alpha + two) * gammaf(beta + one) / gammaf(apb + three);
    fint1 = fint1 * pow(two, apb + two);
    double fint2 = gammaf(alpha + two) * gammaf(beta + two) / gammaf(apb + four);
    fint2 = fint2 * pow(two, apb + three);
    double f2 = (-two * (beta + two) * fint1 + (apb + four) * fint2) * (apb + three) / four;
    if (n == 2) return f2;
    double f3 = 0.;
    for (int i = 3; i <= n; ++i) {
        double di = (double)(i - 1);
        double abn = alpha + beta + di;
        double abnn = abn + di;
        double a1 = -(two * (di + alpha) * (di + beta)) / (abn * abnn * (abnn + one));
        double a2 = (two * (alpha - beta)) / (abnn * (abnn + two));
        double a3 = (two * (abn + one)) / ((abnn + two) * (abnn + one));
        f3 = -(a2 * f2 + a1 * f1) / a3;
        f1 = f2;
        f2 = f3;
    }
    return f3;
}

/* ENDW2, src/speclib.f:327-369 */
static double endw2(int n, double alpha, double beta)
{
    const double zero = 0., one = 1., two = 2., three = 3., four = 4.;
    double apb = alpha + beta;
    if (n == 0) return zero;
    double f1 = gammaf(alpha + one) * gammaf(beta + two) / gammaf(apb + three);
    f1 = f1 * (apb + two) * pow(two, apb + two) / two;
    if (n == 1) return f1;
    double fint1 = gammaf(alpha + one) * gammaf(beta + two) / gammaf(apb + three);
    fint1 = fint1 * pow(two, apb + two);
    double fint2 = gammaf(alpha + two) * gammaf(beta + two) / gammaf(apb + four);
    fint2 = fint2 * pow(two, apb + three);
    double f2 = (two * (alpha + two) * fint1 - (apb + four) * fint2) * (apb + three) / four;
    if (n == 2) return f2;
    double f3 = 0.;
    for (int i = 3; i <= n; ++i) {
        double di = (double)(i - 1);
        double abn = alpha + beta + di;
        double abnn = abn + di;
        double a1 = -(two * (di + alpha) * (di + beta)) / (abn * abnn * (abnn + one));
        double a2 = (two * (alpha - beta)) / (abnn * (abnn + two));
        double a3 = (two * (abn + one)) / ((abnn + two) * (abnn + one));
        f3 = -(a2 * f2 + a1 * f1) / a3;
        f1 = f2;
        f2 = f3;
    }
    return f3;
}

/* ZWGLJD, src/speclib.f:238-281 */
static void zwgljd(double *z, double *w, int np, double alpha, double beta)
{
    int n = np - 1, nm1 = n - 1;
    const double one = 1., two = 2.;
    double p, pd, pm1, pdm1, pm2, pdm2;
    if (nm1 > 0) zwgjd(z + 1, w + 1, nm1, alpha + one, beta + one);
    z[0] = -one;
    z[np - 1] = one;
    for (int i = 1; i < np - 1; ++i) w[i] = w[i] / (one - z[i] * z[i]);
    jacobf(&p, &pd, &pm1, &pdm1, &pm2, &pdm2, n, alpha, beta, z[0]);
    w[0] = endw1(n, alpha, beta) / (two * pd);
    jacobf(&p, &pd, &pm1, &pdm1, &pm2, &pdm2, n, alpha, beta, z[np - 1]);
    w[np - 1] = endw2(n, alpha, beta) / (two * pd);
}

/* ZWGLL, src/speclib.f:107-122 */
void orc_zwgll(double *z, double *w, int np)
{
    zwgljd(z, w, np, 0., 0.);
}

/* fd_weights_full, src/semhat.f:39-93; c(0:n,0:m) column-major */
void orc_fd_weights_full(double xx, const double *x, int n, int m, double *c)
{
#define C(j, k) c[(j) + (n + 1) * (k)]
    double c1 = 1.;
    double c4 = x[0] - xx;
    for (int k = 0; k <= m; ++k)
        for (int j = 0; j <= n; ++j) C(j, k) = 0.;
    C(0, 0) = 1.;
    for (int i = 1; i <= n; ++i) {
        int mn = i < m ? i : m;
        double c2 = 1.;
        double c5 = c4;
        c4 = x[i] - xx;
        for (int j = 0; j <= i - 1; ++j) {
            double c3 = x[i] - x[j];
            c2 = c2 * c3;
            for (int k = mn; k >= 1; --k)
                C(i, k) = c1 * ((double)k * C(i - 1, k - 1) - c5 * C(i - 1, k)) / c2;
            C(i, 0) = -c1 * c5 * C(i - 1, 0) / c2;
            for (int k = mn; k >= 1; --k)
                C(j, k) = (c4 * C(j, k) - (double)k * C(j, k - 1)) / c3;
            C(j, 0) = c4 * C(j, 0) / c3;
        }
        c1 = c2;
    }
#undef C
}

/* semhat, src/semhat.f:2-37 (b = GLL weights, z = GLL points) */
void orc_semhat(double *a, double *b, double *c, double *d, double *z, int n)
{
    int np = n + 1;
    double wk[2 * MAXNX + 2];
    orc_zwgll(z, b, np);
    for (int i = 0; i <= n; ++i) {
        orc_fd_weights_full(z[i], z, n, 1, wk);
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

/* cubic, src/driver.f:196-247 (REAL is real*8 under -fdefault-real-8) */
void orc_cubic(int *mx, int *my, int *mz, int np)
{
    *mx = np;
    *my = 1;
    *mz = 1;
    double ratio = (double)np;
    int iroot3 = (int)(pow((double)np, 1. / 3.) + 0.000001);
    int ix = 0, iy = 0, iz;
    for (int i = iroot3; i >= 1; --i) {
        iz = i;
        int myx = np / iz;
        int nrem = np - myx * iz;
        if (nrem == 0) {
            int iroot2 = (int)(pow((double)myx, 1. / 2.) + 0.000001);
            for (int j = iroot2; j >= 1; --j) {
                iy = j;
                ix = myx / iy;
                nrem = myx - ix * iy;
                if (nrem == 0) break;
            }
            int it;
            if (ix < iy) { it = ix; ix = iy; iy = it; }
            if (ix < iz) { it = ix; ix = iz; iz = it; }
            if (iy < iz) { it = iy; iy = iz; iz = it; }
            if ((double)ix / iz < ratio) {
                ratio = (double)ix / iz;
                *mx = ix;
                *my = iy;
                *mz = iz;
            }
        }
    }
}

/* mod1, src/math.f:418-434 */
static int mod1(int i, int n)
{
    if (i == 0 || n == 0) return 0;
    return (i + n - 1) % n + 1;
}

/* get_exyz, src/prox_dssum.f:75-86 (1-based ex,ey,ez) */
static void get_exyz(int *ex, int *ey, int *ez, int eg, int nelx, int nely)
{
    int nelxy = nelx * nely;
    *ez = 1 + (eg - 1) / nelxy;
    *ey = mod1(eg, nelxy);
    *ey = 1 + (*ey - 1) / nelx;
    *ex = mod1(eg, nelx);
}

/* set_vert_box, src/prox_dssum.f:42-73; lglel(e)=e for one rank
 * (src/driver.f:179-190 with offs=0). */
static void set_vert_box(orc_problem *P)
{
    int nx = P->nx1, nn = nx - 1;
    for (int e = 0; e < P->nelt; ++e) {
        int ex, ey, ez;
        get_exyz(&ex, &ey, &ez, e + 1, P->nelx, P->nely);
        for (int k = 0; k <= nn; ++k)
            for (int j = 0; j <= nn; ++j)
                for (int i = 0; i <= nn; ++i) {
                    int64_t kg = (int64_t)nn * (ez - 1) + k;
                    int64_t jg = (int64_t)nn * (ey - 1) + j;
                    int64_t ig = (int64_t)nn * (ex - 1) + i;
                    int64_t ii = 1 + ig + jg * ((int64_t)nn * P->nelx + 1) +
                                 kg * ((int64_t)nn * P->nelx + 1) * ((int64_t)nn * P->nely + 1);
                    int64_t ll = i + (int64_t)nx * j + (int64_t)nx * nx * k +
                                 (int64_t)nx * nx * nx * e;
                    P->glo[ll] = ii;
                }
    }
}

/* Local gather-scatter map, src/jl/gs.c:88-113 (nonzero_ids: stable sort
 * by id, primary = first index of the id) and :330-357 (local_map: only
 * ids with >= 2 local copies, groups ordered by primary, members in
 * ascending local index).  Bucketing by id in ascending-index order gives
 * exactly that order. */
static void build_gs_map(orc_problem *P)
{
    int64_t n = P->n, maxid = 0;
    for (int64_t l = 0; l < n; ++l)
        if (P->glo[l] > maxid) maxid = P->glo[l];
    int64_t *cnt = calloc((size_t)maxid + 2, sizeof *cnt);
    for (int64_t l = 0; l < n; ++l)
        if (P->glo[l] > 0) cnt[P->glo[l] + 1]++;
    for (int64_t i = 1; i <= maxid + 1; ++i) cnt[i] += cnt[i - 1];
    int64_t *pos = malloc(sizeof *pos * ((size_t)maxid + 1));
    memcpy(pos, cnt, sizeof *pos * ((size_t)maxid + 1));
    int32_t *byid = malloc(sizeof *byid * (size_t)(n ? n : 1));
    for (int64_t l = 0; l < n; ++l)
        if (P->glo[l] > 0) byid[pos[P->glo[l]]++] = (int32_t)l;
    int64_t ng = 0, ns = 0;
    for (int64_t l = 0; l < n; ++l) {
        int64_t id = P->glo[l];
        if (id <= 0) continue;
        int64_t c0 = cnt[id], c1 = cnt[id + 1];
        if (c1 - c0 >= 2 && byid[c0] == l) {
            ng++;
            ns += c1 - c0;
        }
    }
    P->ngroups = ng;
    P->nshared = ns;
    P->goff = malloc(sizeof *P->goff * (size_t)(ng + 1));
    P->gidx = malloc(sizeof *P->gidx * (size_t)(ns ? ns : 1));
    int64_t g = 0, s = 0;
    P->goff[0] = 0;
    for (int64_t l = 0; l < n; ++l) {
        int64_t id = P->glo[l];
        if (id <= 0) continue;
        int64_t c0 = cnt[id], c1 = cnt[id + 1];
        if (c1 - c0 >= 2 && byid[c0] == l) {
            for (int64_t q = c0; q < c1; ++q) P->gidx[s++] = byid[q];
            P->goff[++g] = s;
        }
    }
    free(cnt);
    free(pos);
    free(byid);
}

/* gs_op(...,add) on one rank: gather then scatter,
 * src/jl/gs_local.c:60-87 via src/jl/gs.c:1029-1045. */
void orc_dssum(const orc_problem *P, double *u)
{
#pragma omp parallel for schedule(static) if (P->ngroups > 50000)
    for (int64_t g = 0; g < P->ngroups; ++g) {
        int64_t a = P->goff[g], b = P->goff[g + 1];
        double t = u[P->gidx[a]];
        for (int64_t q = a + 1; q < b; ++q) t += u[P->gidx[q]];
        for (int64_t q = a; q < b; ++q) u[P->gidx[q]] = t;
    }
}

/* get_face, src/cg.f:302-368: which faces of element e (0-based) lie on
 * the domain boundary.  faces[0..5] = x0 (Face4), x1 (Face2), y0 (Face1),
 * y1 (Face3), z0 (Face5), z1 (Face6). */
static void boundary_faces(const orc_problem *P, int e, int faces[6])
{
    int nelx = P->nelx, nely = P->nely, nelz = P->nelz;
    int nelxy = nelx * nely;
    int ngl = e + 1;
    int ir = 1 + (ngl - 1) / nelxy;
    int iq = mod1(ngl, nelxy);
    iq = 1 + (iq - 1) / nelx;
    int ip = mod1(ngl, nelx);
    faces[0] = (ip % nelx == 1) || (nelx == 1);
    faces[1] = (ip % nelx == 0);
    faces[2] = (iq % nely == 1) || (nely == 1);
    faces[3] = (iq % nely == 0);
    faces[4] = (ir % nelz == 1) || (nelz == 1);
    faces[5] = (ir % nelz == 0);
}

static int on_boundary(const int f[6], int i, int j, int k, int nx)
{
    return (f[0] && i == 0) || (f[1] && i == nx - 1) || (f[2] && j == 0) ||
           (f[3] && j == nx - 1) || (f[4] && k == 0) || (f[5] && k == nx - 1);
}

/* Full list: every point zeroed by get_face on any element.  Ref list:
 * the element-1 positions that src/cg.f:245-249 records (nelt copies of
 * the same positions, de-duplicated here; zeroing is idempotent). */
static void build_masks(orc_problem *P)
{
    int nx = P->nx1;
    int64_t nxyz = (int64_t)nx * nx * nx;
    int64_t cnt = 0;
    for (int e = 0; e < P->nelt; ++e) {
        int f[6];
        boundary_faces(P, e, f);
        for (int k = 0; k < nx; ++k)
            for (int j = 0; j < nx; ++j)
                for (int i = 0; i < nx; ++i) cnt += on_boundary(f, i, j, k, nx);
    }
    P->nfull = cnt;
    P->mfull = malloc(sizeof *P->mfull * (size_t)(cnt ? cnt : 1));
    cnt = 0;
    for (int e = 0; e < P->nelt; ++e) {
        int f[6];
        boundary_faces(P, e, f);
        for (int k = 0; k < nx; ++k)
            for (int j = 0; j < nx; ++j)
                for (int i = 0; i < nx; ++i)
                    if (on_boundary(f, i, j, k, nx))
                        P->mfull[cnt++] = (int32_t)(i + nx * j + nx * nx * k + nxyz * e);
    }
    int f1[6];
    boundary_faces(P, 0, f1);
    cnt = 0;
    for (int64_t l = 0; l < nxyz; ++l) {
        int i = (int)(l % nx), j = (int)((l / nx) % nx), k = (int)(l / (nx * nx));
        cnt += on_boundary(f1, i, j, k, nx);
    }
    P->nref = cnt;
    P->mref = malloc(sizeof *P->mref * (size_t)(cnt ? cnt : 1));
    cnt = 0;
    for (int64_t l = 0; l < nxyz; ++l) {
        int i = (int)(l % nx), j = (int)((l / nx) % nx), k = (int)(l / (nx * nx));
        if (on_boundary(f1, i, j, k, nx)) P->mref[cnt++] = (int32_t)l;
    }
}

/* setup_g, src/prox_setup.f:27-52 */
static void setup_g_brick(orc_problem *P)
{
    int nx = P->nx1;
    int64_t l = 0;
    for (int e = 0; e < P->nelt; ++e)
        for (int k = 0; k < nx; ++k)
            for (int j = 0; j < nx; ++j)
                for (int i = 0; i < nx; ++i, ++l) {
                    double *g = P->g + 6 * l;
                    double v = P->w[i] * P->w[j] * P->w[k];
                    g[0] = v;
                    g[1] = 0.;
                    g[2] = 0.;
                    g[3] = v;
                    g[4] = 0.;
                    g[5] = v;
                }
}

orc_problem *orc_create(int nelx, int nely, int nelz, int nx1)
{
    if (nx1 < 2 || nx1 > MAXNX || nelx < 1 || nely < 1 || nelz < 1) return NULL;
    orc_problem *P = calloc(1, sizeof *P);
    P->nelx = nelx;
    P->nely = nely;
    P->nelz = nelz;
    P->nelt = nelx * nely * nelz;
    P->nx1 = nx1;
    P->n = (int64_t)nx1 * nx1 * nx1 * P->nelt;
    P->glo = malloc(sizeof *P->glo * (size_t)P->n);
    P->c = malloc(sizeof *P->c * (size_t)P->n);
    P->g = malloc(sizeof *P->g * 6 * (size_t)P->n);
    P->wk = malloc(sizeof *P->wk * (size_t)P->n);
    /* proxy_setupds (src/prox_dssum.f:14-40) */
    set_vert_box(P);
    build_gs_map(P);
    /* set_multiplicity (src/driver.f:274-291) */
    for (int64_t l = 0; l < P->n; ++l) P->c[l] = 1.0;
    orc_dssum(P, P->c);
    for (int64_t l = 0; l < P->n; ++l) P->c[l] = 1. / P->c[l];
    /* proxy_setup (src/prox_setup.f:2-25) */
    double a[MAXNX * MAXNX], cc[MAXNX * MAXNX];
    orc_semhat(a, P->w, cc, P->d, P->z, nx1 - 1);
    for (int j = 0; j < nx1; ++j)
        for (int i = 0; i < nx1; ++i) P->dt[i + nx1 * j] = P->d[j + nx1 * i];
    setup_g_brick(P);
    build_masks(P);
    P->mask_mode = ORC_MASK_REF;
    P->mask_built = 0;
    return P;
}

void orc_free(orc_problem *P)
{
    if (!P) return;
    free(P->glo);
    free(P->c);
    free(P->g);
    free(P->wk);
    free(P->goff);
    free(P->gidx);
    free(P->mfull);
    free(P->mref);
    free(P->dinv);
    hsmg_free(P->mg);
    free(P);
}

int64_t orc_npts(const orc_problem *P) { return P->n; }
void orc_get_dxm1(const orc_problem *P, double *d) { memcpy(d, P->d, sizeof(double) * P->nx1 * P->nx1); }
void orc_get_c(const orc_problem *P, double *c) { memcpy(c, P->c, sizeof(double) * (size_t)P->n); }
void orc_get_g(const orc_problem *P, double *g) { memcpy(g, P->g, sizeof(double) * 6 * (size_t)P->n); }
void orc_set_g(orc_problem *P, const double *g) { memcpy(P->g, g, sizeof(double) * 6 * (size_t)P->n); }
void orc_get_glo(const orc_problem *P, int64_t *glo) { memcpy(glo, P->glo, sizeof(int64_t) * (size_t)P->n); }
int64_t orc_gs_ngroups(const orc_problem *P) { return P->ngroups; }
int64_t orc_gs_nshared(const orc_problem *P) { return P->nshared; }
void orc_gs_map(const orc_problem *P, int64_t *off, int32_t *idx)
{
    memcpy(off, P->goff, sizeof(int64_t) * (size_t)(P->ngroups + 1));
    memcpy(idx, P->gidx, sizeof(int32_t) * (size_t)P->nshared);
}
int64_t orc_mask_full_count(const orc_problem *P) { return P->nfull; }
int64_t orc_mask_ref_count(const orc_problem *P) { return P->nref; }
void orc_mask_lists(const orc_problem *P, int32_t *full, int32_t *ref)
{
    memcpy(full, P->mfull, sizeof(int32_t) * (size_t)P->nfull);
    memcpy(ref, P->mref, sizeof(int32_t) * (size_t)P->nref);
}
void orc_set_mask_mode(orc_problem *P, int mode) { P->mask_mode = mode; }
void orc_reset_mask(orc_problem *P) { P->mask_built = 0; }

/* maskit, src/cg.f:214-258.  First call after a reset zeroes every
 * domain-boundary point (get_face on each element) and records the list;
 * later calls zero the recorded list.  In DIRICHLET mode the full list is
 * applied every time. */
void orc_maskit(orc_problem *P, double *w)
{
    if (P->mask_mode == ORC_MASK_DIRICHLET || !P->mask_built) {
        for (int64_t q = 0; q < P->nfull; ++q) w[P->mfull[q]] = 0.0;
        P->mask_built = 1;
    } else {
        for (int64_t q = 0; q < P->nref; ++q) w[P->mref[q]] = 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* ax_e: src/cg.f:142-212 with mxm -> mxfN (src/mxm_std.f:1-66,         */
/* 68-631): c(i,j) = a(i,1)b(1,j) + a(i,2)b(2,j) + ... left to right.   */
/* ------------------------------------------------------------------ */

/* C(n1,n3) = A(n1,n2) B(n2,n3), all column-major and contiguous. */
static void mxm(const double *a, int n1, const double *b, int n2, double *c, int n3)
{
    for (int j = 0; j < n3; ++j)
        for (int i = 0; i < n1; ++i) {
            double s = a[i] * b[n2 * j];
            for (int l = 1; l < n2; ++l) s = s + a[i + n1 * l] * b[l + n2 * j];
            c[i + n1 * j] = s;
        }
}

/* local_grad3, src/cg.f:170-187 */
static void local_grad3(double *ur, double *us, double *ut, const double *u, int n,
                        const double *D, const double *Dt)
{
    int m1 = n + 1, m2 = m1 * m1;
    mxm(D, m1, u, m1, ur, m2);
    for (int k = 0; k <= n; ++k) mxm(u + m2 * k, m1, Dt, m1, us + m2 * k, m1);
    mxm(u, m2, Dt, m1, ut, m1);
}

/* local_grad3_t, src/cg.f:189-212 */
static void local_grad3_t(double *u, const double *ur, const double *us, const double *ut,
                          int N, const double *D, const double *Dt, double *w)
{
    int m1 = N + 1, m2 = m1 * m1, m3 = m2 * m1;
    mxm(Dt, m1, ur, m1, u, m2);
    for (int k = 0; k <= N; ++k) mxm(us + m2 * k, m1, D, m1, w + m2 * k, m1);
    for (int i = 0; i < m3; ++i) u[i] = u[i] + w[i];
    mxm(ut, m2, D, m1, w, m1);
    for (int i = 0; i < m3; ++i) u[i] = u[i] + w[i];
}

/* ax_e, src/cg.f:142-168 */
void orc_ax_e(const orc_problem *P, int e, double *w, const double *u)
{
    int nx = P->nx1, nxyz = nx * nx * nx;
    double ur[MAXNX * MAXNX * MAXNX], us[MAXNX * MAXNX * MAXNX], ut[MAXNX * MAXNX * MAXNX],
        wk[MAXNX * MAXNX * MAXNX];
    const double *g = P->g + 6 * (int64_t)nxyz * e;
    local_grad3(ur, us, ut, u, nx - 1, P->d, P->dt);
    for (int i = 0; i < nxyz; ++i) {
        const double *gi = g + 6 * i;
        double wr = gi[0] * ur[i] + gi[1] * us[i] + gi[2] * ut[i];
        double ws = gi[1] * ur[i] + gi[3] * us[i] + gi[4] * ut[i];
        double wt = gi[2] * ur[i] + gi[4] * us[i] + gi[5] * ut[i];
        ur[i] = wr;
        us[i] = ws;
        ut[i] = wt;
    }
    local_grad3_t(w, ur, us, ut, nx - 1, P->d, P->dt, wk);
}

/* ax, src/cg.f:96-125 */
void orc_ax(orc_problem *P, double *w, const double *u)
{
    int64_t nxyz = (int64_t)P->nx1 * P->nx1 * P->nx1;
#pragma omp parallel for schedule(static) if (P->nelt > 64)
    for (int e = 0; e < P->nelt; ++e) orc_ax_e(P, e, w + nxyz * e, u + nxyz * e);
    orc_dssum(P, w);
#pragma omp parallel for schedule(static) if (P->n > 1000000)
    for (int64_t i = 0; i < P->n; ++i) w[i] = w[i] + 0.1 * u[i]; /* add2s2, math.f:544 */
    orc_maskit(P, w);
}

/* glsc3, src/math.f:603-617 (gop is the identity on one rank) */
double orc_glsc3(const double *a, const double *b, const double *m, int64_t n)
{
    double tmp = 0.0;
    for (int64_t i = 0; i < n; ++i) tmp = tmp + a[i] * b[i] * m[i];
    return tmp;
}

/* Neumaier's compensated sum step (Neumaier 1974, the improved
 * Kahan-Babuska summation): s + c carries the running sum. */
static void neumaier_add(double *s, double *c, double t)
{
    double u = *s + t;
    if (fabs(*s) >= fabs(t))
        *c += (*s - u) + t;
    else
        *c += (t - u) + *s;
    *s = u;
}

/* The sum of glsc3's rounded terms (a_i b_i) m_i, correctly rounded to
 * within a unit in the last place (compensated summation; its error is
 * 2 eps |sum| + n eps^2 sum |t_i|).  Not the reference's value: it takes
 * the summation order out of the comparison (the reference's sequential
 * sum of n terms errs by up to n eps sum |t_i|). */
double orc_glsc3_exact(const double *a, const double *b, const double *m, int64_t n)
{
    enum { NT = 256 };
    double ps[NT] = {0}, pc[NT] = {0};
    int nt = 1;
#pragma omp parallel if (n > 1000000)
    {
#pragma omp single
        nt = omp_get_num_threads() < NT ? omp_get_num_threads() : NT;
        int t = omp_get_thread_num();
        if (t < nt) {
            int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
            double s = 0.0, c = 0.0;
            for (int64_t i = lo; i < hi; ++i) neumaier_add(&s, &c, a[i] * b[i] * m[i]);
            ps[t] = s;
            pc[t] = c;
        }
    }
    double s = 0.0, c = 0.0;
    for (int t = 0; t < nt; ++t) {
        neumaier_add(&s, &c, ps[t]);
        neumaier_add(&s, &c, pc[t]);
    }
    return s + c;
}

void orc_set_exact_dots(orc_problem *P, int on) { P->exact_dots = on ? 1 : 0; }

static double cg_dot(const orc_problem *P, const double *a, const double *b, const double *m,
                     int64_t n)
{
    return P->exact_dots ? orc_glsc3_exact(a, b, m, n) : orc_glsc3(a, b, m, n);
}

/* ------------------------------------------------------------------ */
/* Jacobi (not in the reference).  Local diagonal of A_L = sum_ab       */
/* D_a^T G_ab D_b at point (i,j,k):                                      */
/*   t1 = sum_m (D(m,i)*D(m,i))*g1(m,j,k)   m = 0..n in order            */
/*   t2 = sum_m (D(m,j)*D(m,j))*g4(i,m,k)                                */
/*   t3 = sum_m (D(m,k)*D(m,k))*g6(i,j,m)                                */
/*   x2 = (2*(D(i,i)*D(j,j)))*g2, x3 = (2*(D(i,i)*D(k,k)))*g3,           */
/*   x5 = (2*(D(j,j)*D(k,k)))*g5                                         */
/*   d  = ((((t1+t2)+t3)+x2)+x3)+x5                                      */
/* ------------------------------------------------------------------ */
void orc_jacobi_diag_local(const orc_problem *P, double *dl)
{
    int nx = P->nx1;
    int64_t nxyz = (int64_t)nx * nx * nx;
    const double *D = P->d;
#pragma omp parallel for schedule(static) if (P->nelt > 64)
    for (int e = 0; e < P->nelt; ++e) {
        const double *g = P->g + 6 * nxyz * e;
        for (int k = 0; k < nx; ++k)
            for (int j = 0; j < nx; ++j)
                for (int i = 0; i < nx; ++i) {
                    int p = i + nx * j + nx * nx * k;
                    double t1 = (D[0 + nx * i] * D[0 + nx * i]) * g[6 * (0 + nx * j + nx * nx * k) + 0];
                    double t2 = (D[0 + nx * j] * D[0 + nx * j]) * g[6 * (i + nx * 0 + nx * nx * k) + 3];
                    double t3 = (D[0 + nx * k] * D[0 + nx * k]) * g[6 * (i + nx * j + nx * nx * 0) + 5];
                    for (int m = 1; m < nx; ++m) {
                        t1 = t1 + (D[m + nx * i] * D[m + nx * i]) * g[6 * (m + nx * j + nx * nx * k) + 0];
                        t2 = t2 + (D[m + nx * j] * D[m + nx * j]) * g[6 * (i + nx * m + nx * nx * k) + 3];
                        t3 = t3 + (D[m + nx * k] * D[m + nx * k]) * g[6 * (i + nx * j + nx * nx * m) + 5];
                    }
                    double dii = D[i + nx * i], djj = D[j + nx * j], dkk = D[k + nx * k];
                    double x2 = (2.0 * (dii * djj)) * g[6 * p + 1];
                    double x3 = (2.0 * (dii * dkk)) * g[6 * p + 2];
                    double x5 = (2.0 * (djj * dkk)) * g[6 * p + 4];
                    dl[nxyz * e + p] = ((((t1 + t2) + t3) + x2) + x3) + x5;
                }
    }
}

void orc_setup_jacobi(orc_problem *P)
{
    if (!P->dinv) P->dinv = malloc(sizeof *P->dinv * (size_t)P->n);
    orc_jacobi_diag_local(P, P->dinv);
    orc_dssum(P, P->dinv);
    for (int64_t i = 0; i < P->n; ++i) P->dinv[i] = 1.0 / (P->dinv[i] + 0.1);
}

void orc_get_dinv(const orc_problem *P, double *dinv)
{
    memcpy(dinv, P->dinv, sizeof(double) * (size_t)P->n);
}

/* ------------------------------------------------------------------ */
/* cg, src/cg.f:2-84                                                     */
/* ------------------------------------------------------------------ */
void orc_cg(orc_problem *P, double *x, const double *f, double *r, int niter, int precond,
            double *hist)
{
    int64_t n = P->n;
    double *w = malloc(sizeof *w * (size_t)n);
    double *p = calloc((size_t)n, sizeof *p);
    double *z = malloc(sizeof *z * (size_t)n);
    const double *c = P->c;
    double rtz1 = 1.0, rtz2, beta, alpha, alphm, pap, rtr;
    int h = 0;

#pragma omp parallel for schedule(static) if (n > 1000000)
    for (int64_t i = 0; i < n; ++i) {
        x[i] = 0.0;
        r[i] = f[i];
    }
    orc_maskit(P, r);
    hist[h++] = cg_dot(P, r, c, r, n);
    for (int iter = 1; iter <= niter; ++iter) {
        /* solveM (src/cg.f:86-94): identity copy (src/hsmg_dum.f:9-15)
         * or Jacobi z = r * dinv. */
        if (precond == ORC_PRECOND_JACOBI) {
#pragma omp parallel for schedule(static) if (n > 1000000)
            for (int64_t i = 0; i < n; ++i) z[i] = r[i] * P->dinv[i];
        } else if (precond == ORC_PRECOND_HSMG) {
            orc_hsmg_solve(P, z, r); /* h1mg_solve masks r in place */
        } else {
#pragma omp parallel for schedule(static) if (n > 1000000)
            for (int64_t i = 0; i < n; ++i) z[i] = r[i];
        }
        rtz2 = rtz1;
        rtz1 = cg_dot(P, r, c, z, n);
        hist[h++] = rtz1;
        beta = rtz1 / rtz2;
        if (iter == 1) beta = 0.0;
#pragma omp parallel for schedule(static) if (n > 1000000)
        for (int64_t i = 0; i < n; ++i) p[i] = beta * p[i] + z[i]; /* add2s1 */
        orc_ax(P, w, p);
        pap = cg_dot(P, w, c, p, n);
        hist[h++] = pap;
        alpha = rtz1 / pap;
        alphm = -alpha;
#pragma omp parallel for schedule(static) if (n > 1000000)
        for (int64_t i = 0; i < n; ++i) {
            x[i] = x[i] + alpha * p[i]; /* add2s2 */
            r[i] = r[i] + alphm * w[i];
        }
        rtr = cg_dot(P, r, c, r, n);
        hist[h++] = rtr;
    }
    free(w);
    free(p);
    free(z);
}

/* ------------------------------------------------------------------ */
/* Synthetic inputs                                                      */
/* ------------------------------------------------------------------ */

/* Counter-based splitmix64: value k = mix(seed + (k+1)*golden). */
static uint64_t splitmix64_at(uint64_t seed, uint64_t k)
{
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void orc_rand_uniform(double *f, int64_t n, uint64_t seed)
{
    for (int64_t i = 0; i < n; ++i) {
        double u = (double)(splitmix64_at(seed, (uint64_t)i) >> 11) * 0x1.0p-53;
        f[i] = 2.0 * u - 1.0;
    }
}

/* set_f's post-processing, src/driver.f:96-97: dssum then col2 by c. */
void orc_rhs_from_raw(const orc_problem *P, double *f)
{
    orc_dssum(P, f);
    for (int64_t i = 0; i < P->n; ++i) f[i] = f[i] * P->c[i];
}

/* ------------------------------------------------------------------ */
/* Deformed geometry (not in the reference; prox_setup.f:27-52 is the   */
/* brick case it must reduce to).                                        */
/* ------------------------------------------------------------------ */
static double lerp(double a, double b, double s)
{
    return a + (b - a) * ((1.0 + s) * 0.5);
}

void orc_deformed_g(orc_problem *P, double amp, uint64_t seed)
{
    int nx = P->nx1;
    int vx = P->nelx + 1, vy = P->nely + 1, vz = P->nelz + 1;
    int64_t nv = (int64_t)vx * vy * vz;
    double *V = malloc(sizeof *V * 3 * (size_t)nv);
    orc_rand_uniform(V, 3 * nv, seed);
    double scale = amp * 2.0;
    for (int c = 0; c < vz; ++c)
        for (int b = 0; b < vy; ++b)
            for (int a = 0; a < vx; ++a) {
                int64_t v = a + (int64_t)vx * (b + (int64_t)vy * c);
                V[3 * v + 0] = 2.0 * a + scale * V[3 * v + 0];
                V[3 * v + 1] = 2.0 * b + scale * V[3 * v + 1];
                V[3 * v + 2] = 2.0 * c + scale * V[3 * v + 2];
            }
    int64_t nxyz = (int64_t)nx * nx * nx;
    for (int e = 0; e < P->nelt; ++e) {
        int ex, ey, ez;
        get_exyz(&ex, &ey, &ez, e + 1, P->nelx, P->nely);
        ex -= 1;
        ey -= 1;
        ez -= 1;
        double X[2][2][2][3]; /* X[c][b][a][d] corner (a,b,c) */
        for (int cc = 0; cc < 2; ++cc)
            for (int bb = 0; bb < 2; ++bb)
                for (int aa = 0; aa < 2; ++aa) {
                    int64_t v = (ex + aa) + (int64_t)vx * ((ey + bb) + (int64_t)vy * (ez + cc));
                    for (int d = 0; d < 3; ++d) X[cc][bb][aa][d] = V[3 * v + d];
                }
        for (int k = 0; k < nx; ++k)
            for (int j = 0; j < nx; ++j)
                for (int i = 0; i < nx; ++i) {
                    double r = P->z[i], s = P->z[j], t = P->z[k];
                    double J[3][3]; /* J[d][a] = dx_d / dr_a */
                    for (int d = 0; d < 3; ++d) {
                        /* d/dr: edge differences along a, bilinear in (s,t) */
                        double e00 = X[0][0][1][d] - X[0][0][0][d];
                        double e10 = X[0][1][1][d] - X[0][1][0][d];
                        double e01 = X[1][0][1][d] - X[1][0][0][d];
                        double e11 = X[1][1][1][d] - X[1][1][0][d];
                        J[d][0] = 0.5 * lerp(lerp(e00, e10, s), lerp(e01, e11, s), t);
                        /* d/ds: differences along b, bilinear in (r,t) */
                        double f00 = X[0][1][0][d] - X[0][0][0][d];
                        double f10 = X[0][1][1][d] - X[0][0][1][d];
                        double f01 = X[1][1][0][d] - X[1][0][0][d];
                        double f11 = X[1][1][1][d] - X[1][0][1][d];
                        J[d][1] = 0.5 * lerp(lerp(f00, f10, r), lerp(f01, f11, r), t);
                        /* d/dt: differences along c, bilinear in (r,s) */
                        double h00 = X[1][0][0][d] - X[0][0][0][d];
                        double h10 = X[1][0][1][d] - X[0][0][1][d];
                        double h01 = X[1][1][0][d] - X[0][1][0][d];
                        double h11 = X[1][1][1][d] - X[0][1][1][d];
                        J[d][2] = 0.5 * lerp(lerp(h00, h10, r), lerp(h01, h11, r), s);
                    }
                    double xr = J[0][0], xs = J[0][1], xt = J[0][2];
                    double yr = J[1][0], ys = J[1][1], yt = J[1][2];
                    double zr = J[2][0], zs = J[2][1], zt = J[2][2];
                    /* cofactors C[d][a] of J[d][a], written without negation */
                    double C00 = ys * zt - yt * zs, C10 = xt * zs - xs * zt, C20 = xs * yt - xt * ys;
                    double C01 = yt * zr - yr * zt, C11 = xr * zt - xt * zr, C21 = xt * yr - xr * yt;
                    double C02 = yr * zs - ys * zr, C12 = xs * zr - xr * zs, C22 = xr * ys - xs * yr;
                    double det = (xr * C00 + xs * C01) + xt * C02;
                    double w3 = P->w[i] * P->w[j] * P->w[k];
                    double Grr = (C00 * C00 + C10 * C10) + C20 * C20;
                    double Grs = (C00 * C01 + C10 * C11) + C20 * C21;
                    double Grt = (C00 * C02 + C10 * C12) + C20 * C22;
                    double Gss = (C01 * C01 + C11 * C11) + C21 * C21;
                    double Gst = (C01 * C02 + C11 * C12) + C21 * C22;
                    double Gtt = (C02 * C02 + C12 * C12) + C22 * C22;
                    double *g = P->g + 6 * (nxyz * e + i + nx * j + nx * nx * k);
                    g[0] = (w3 * Grr) / det;
                    g[1] = (w3 * Grs) / det;
                    g[2] = (w3 * Grt) / det;
                    g[3] = (w3 * Gss) / det;
                    g[4] = (w3 * Gst) / det;
                    g[5] = (w3 * Gtt) / det;
                }
    }
    free(V);
}

/* ================================================================== */
/* HSMG preconditioner: src/hsmg.f                                      */
/* ================================================================== */

typedef struct {
    int64_t ng;
    int64_t *off;
    int32_t *idx;
} gsmap;

typedef struct {
    int nx, nh, nl; /* polynomial order, points per direction, nh+2 */
    int64_t n;      /* nh^3 * nelt */
    double ah[MAXNX * MAXNX], bh[MAXNX], ch[MAXNX * MAXNX], dh[MAXNX * MAXNX], zh[MAXNX];
    double jh[MAXNX * MAXNX], jht[MAXNX * MAXNX]; /* to level l+1: jh(nf,nc), jht(nc,nf) */
    gsmap gs, gss;  /* nh and nh+2 numberings */
    double *rwt;    /* restriction weights wt(nh,nh,2,3,nelt) */
    double *fs;     /* FDM s(nl*nl,2,3,nelt) */
    double *fd;     /* FDM d(nl^3,nelt) */
    double *swt;    /* Schwarz weights wt(nh,nh,4,3,nelt) */
    int32_t *mask;
    int64_t nmask;
} mglevel;

struct orc_hsmg {
    int lmax;
    mglevel L[4]; /* 1..lmax */
    double *e, *w, *r, *work;
    /* optional coarse solve on level 1 (orc_hsmg_set_coarse) */
    int nc;
    double *cainv;  /* nc x nc, column-major */
    int32_t *ccidx; /* level-1 point -> unknown, -1 masked */
};

static void gs_from_glo(const int64_t *glo, int64_t n, gsmap *m)
{
    int64_t maxid = 0;
    for (int64_t l = 0; l < n; ++l)
        if (glo[l] > maxid) maxid = glo[l];
    int64_t *cnt = calloc((size_t)maxid + 2, sizeof *cnt);
    for (int64_t l = 0; l < n; ++l) cnt[glo[l] + 1]++;
    for (int64_t i = 1; i <= maxid + 1; ++i) cnt[i] += cnt[i - 1];
    int64_t *pos = malloc(sizeof *pos * ((size_t)maxid + 1));
    memcpy(pos, cnt, sizeof *pos * ((size_t)maxid + 1));
    int32_t *byid = malloc(sizeof *byid * (size_t)(n ? n : 1));
    for (int64_t l = 0; l < n; ++l) byid[pos[glo[l]]++] = (int32_t)l;
    int64_t ng = 0, ns = 0;
    for (int64_t l = 0; l < n; ++l) {
        int64_t c0 = cnt[glo[l]], c1 = cnt[glo[l] + 1];
        if (c1 - c0 >= 2 && byid[c0] == l) {
            ng++;
            ns += c1 - c0;
        }
    }
    m->ng = ng;
    m->off = malloc(sizeof *m->off * (size_t)(ng + 1));
    m->idx = malloc(sizeof *m->idx * (size_t)(ns ? ns : 1));
    int64_t g = 0, s = 0;
    m->off[0] = 0;
    for (int64_t l = 0; l < n; ++l) {
        int64_t c0 = cnt[glo[l]], c1 = cnt[glo[l] + 1];
        if (c1 - c0 >= 2 && byid[c0] == l) {
            for (int64_t q = c0; q < c1; ++q) m->idx[s++] = byid[q];
            m->off[++g] = s;
        }
    }
    free(cnt);
    free(pos);
    free(byid);
}

static void gs_apply(const gsmap *m, double *u)
{
    for (int64_t g = 0; g < m->ng; ++g) {
        int64_t a = m->off[g], b = m->off[g + 1];
        double t = u[m->idx[a]];
        for (int64_t q = a + 1; q < b; ++q) t += u[m->idx[q]];
        for (int64_t q = a; q < b; ++q) u[m->idx[q]] = t;
    }
}

/* set_vert_box (src/prox_dssum.f:42-73) for nx points per direction. */
static int64_t *vert_box_nx(const orc_problem *P, int nx)
{
    int nn = nx - 1;
    int64_t nxyz = (int64_t)nx * nx * nx;
    int64_t *glo = malloc(sizeof *glo * (size_t)(nxyz * P->nelt));
    for (int e = 0; e < P->nelt; ++e) {
        int ex, ey, ez;
        get_exyz(&ex, &ey, &ez, e + 1, P->nelx, P->nely);
        for (int k = 0; k <= nn; ++k)
            for (int j = 0; j <= nn; ++j)
                for (int i = 0; i <= nn; ++i) {
                    int64_t kg = (int64_t)nn * (ez - 1) + k;
                    int64_t jg = (int64_t)nn * (ey - 1) + j;
                    int64_t ig = (int64_t)nn * (ex - 1) + i;
                    glo[i + nx * j + (int64_t)nx * nx * k + nxyz * e] =
                        1 + ig + jg * ((int64_t)nn * P->nelx + 1) +
                        kg * ((int64_t)nn * P->nelx + 1) * ((int64_t)nn * P->nely + 1);
                }
    }
    return glo;
}

/* Generalized symmetric-definite eigenproblem A x = lam B x
 * (generalev, src/hsmg.f:483-523, which calls LAPACK DSYGV with
 * itype=1, jobz='V', uplo='U').  Same definition of the result: lam
 * ascending, eigenvectors in the columns of a with X^T B X = I.  Method:
 * Cholesky B = U^T U, C = U^-T A U^-1, cyclic Jacobi rotations on C
 * (Golub & Van Loan, Matrix Computations, 4th ed., sec. 8.5) until the
 * off-diagonal part is below 1e-30 of the Frobenius norm, X = U^-1 Y.
 * Eigenvector signs may differ from LAPACK's; every use below is of the
 * form S diag S^T, which does not depend on them. */
static void generalev(double *a, const double *b, double *lam, int n)
{
#define A_(i, j) a[(i) + n * (j)]
    double U[MAXNX * MAXNX] = {0}, C[MAXNX * MAXNX], T[MAXNX * MAXNX], Y[MAXNX * MAXNX];
    /* Cholesky, upper: B = U^T U */
    for (int j = 0; j < n; ++j) {
        double s = b[j + n * j];
        for (int k = 0; k < j; ++k) s -= U[k + n * j] * U[k + n * j];
        U[j + n * j] = sqrt(s);
        for (int i = j + 1; i < n; ++i) {
            double t = b[j + n * i];
            for (int k = 0; k < j; ++k) t -= U[k + n * j] * U[k + n * i];
            U[j + n * i] = t / U[j + n * j];
        }
    }
    /* T = A U^-1 (solve T U = A row by row), C = U^-T T */
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double t = A_(i, j);
            for (int k = 0; k < j; ++k) t -= T[i + n * k] * U[k + n * j];
            T[i + n * j] = t / U[j + n * j];
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double t = T[i + n * j];
            for (int k = 0; k < i; ++k) t -= U[k + n * i] * C[k + n * j];
            C[i + n * j] = t / U[i + n * i];
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i) {
            double m = 0.5 * (C[i + n * j] + C[j + n * i]);
            C[i + n * j] = C[j + n * i] = m;
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) Y[i + n * j] = (i == j) ? 1.0 : 0.0;
    double fro = 0.0;
    for (int i = 0; i < n * n; ++i) fro += C[i] * C[i];
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < j; ++i) off += C[i + n * j] * C[i + n * j];
        if (off <= 1e-30 * fro) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = C[p + n * q];
                if (apq == 0.0) continue;
                double app = C[p + n * p], aqq = C[q + n * q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double ckp = C[k + n * p], ckq = C[k + n * q];
                    C[k + n * p] = c * ckp - s * ckq;
                    C[k + n * q] = s * ckp + c * ckq;
                }
                for (int k = 0; k < n; ++k) {
                    double cpk = C[p + n * k], cqk = C[q + n * k];
                    C[p + n * k] = c * cpk - s * cqk;
                    C[q + n * k] = s * cpk + c * cqk;
                }
                for (int k = 0; k < n; ++k) {
                    double ykp = Y[k + n * p], ykq = Y[k + n * q];
                    Y[k + n * p] = c * ykp - s * ykq;
                    Y[k + n * q] = s * ykp + c * ykq;
                }
            }
    }
    int ord[MAXNX];
    for (int i = 0; i < n; ++i) ord[i] = i;
    for (int i = 1; i < n; ++i) /* stable insertion sort by eigenvalue */
        for (int k = i; k > 0 && C[ord[k] * (n + 1)] < C[ord[k - 1] * (n + 1)]; --k) {
            int t = ord[k];
            ord[k] = ord[k - 1];
            ord[k - 1] = t;
        }
    for (int jj = 0; jj < n; ++jj) {
        int j = ord[jj];
        lam[jj] = C[j * (n + 1)];
        /* x = U^-1 y (back substitution) */
        for (int i = n - 1; i >= 0; --i) {
            double t = Y[i + n * j];
            for (int k = i + 1; k < n; ++k) t -= U[i + n * k] * A_(k, jj);
            A_(i, jj) = t / U[i + n * i];
        }
    }
#undef A_
}

/* h1mg_setup_fast1d_a / _b (src/hsmg.f:402-481): the 1-D stiffness and
 * mass of the element extended by one point of each neighbour; a and b
 * are (n+3)x(n+3), ah (n+1)x(n+1), bh (n+1). */
static void fast1d_ab(double *a, double *b, int lbc, int rbc, double ll, double lm, double lr,
                      const double *ah, const double *bh, int n)
{
    const int m = n + 3;
#define A_(i, j) a[(i) + m * (j)]
#define B_(i, j) b[(i) + m * (j)]
#define AH(i, j) ah[(i) + (n + 1) * (j)]
    int i0 = (lbc == 1) ? 1 : 0, i1 = (rbc == 1) ? n - 1 : n;
    for (int i = 0; i < m * m; ++i) a[i] = b[i] = 0.0;
    double fac = 2.0 / lm;
    A_(1, 1) = 1.0;
    A_(n + 1, n + 1) = 1.0;
    for (int j = i0; j <= i1; ++j)
        for (int i = i0; i <= i1; ++i) A_(i + 1, j + 1) = fac * AH(i, j);
    if (lbc == 0) {
        fac = 2.0 / ll;
        A_(0, 0) = fac * AH(n - 1, n - 1);
        A_(1, 0) = fac * AH(n, n - 1);
        A_(0, 1) = fac * AH(n - 1, n);
        A_(1, 1) = A_(1, 1) + fac * AH(n, n);
    } else {
        A_(0, 0) = 1.0;
    }
    if (rbc == 0) {
        fac = 2.0 / lr;
        A_(n + 1, n + 1) = A_(n + 1, n + 1) + fac * AH(0, 0);
        A_(n + 2, n + 1) = fac * AH(1, 0);
        A_(n + 1, n + 2) = fac * AH(0, 1);
        A_(n + 2, n + 2) = fac * AH(1, 1);
    } else {
        A_(n + 2, n + 2) = 1.0;
    }
    fac = 0.5 * lm;
    B_(1, 1) = 1.0;
    B_(n + 1, n + 1) = 1.0;
    for (int i = i0; i <= i1; ++i) B_(i + 1, i + 1) = fac * bh[i];
    if (lbc == 0) {
        fac = 0.5 * ll;
        B_(0, 0) = fac * bh[n - 1];
        B_(1, 1) = B_(1, 1) + fac * bh[n];
    } else {
        B_(0, 0) = 1.0;
    }
    if (rbc == 0) {
        fac = 0.5 * lr;
        B_(n + 1, n + 1) = B_(n + 1, n + 1) + fac * bh[0];
        B_(n + 2, n + 2) = fac * bh[1];
    } else {
        B_(n + 2, n + 2) = 1.0;
    }
#undef A_
#undef B_
#undef AH
}

/* h1mg_setup_fast1d (src/hsmg.f:378-400): s(nl,nl,2) = eigenvectors with
 * the Dirichlet rows zeroed, and their transpose. */
static void fast1d(double *s, double *lam, int nl, int lbc, int rbc, double ll, double lm,
                   double lr, const double *ah, const double *bh, int n)
{
    double b[MAXNX * MAXNX];
    fast1d_ab(s, b, lbc, rbc, ll, lm, lr, ah, bh, n);
    generalev(s, b, lam, nl);
    if (lbc > 0)
        for (int j = 0; j < nl; ++j) s[0 + nl * j] = 0.;
    if (lbc == 1)
        for (int j = 0; j < nl; ++j) s[1 + nl * j] = 0.;
    if (rbc > 0)
        for (int j = 0; j < nl; ++j) s[nl - 1 + nl * j] = 0.;
    if (rbc == 1)
        for (int j = 0; j < nl; ++j) s[nl - 2 + nl * j] = 0.;
    for (int j = 0; j < nl; ++j)
        for (int i = 0; i < nl; ++i) s[nl * nl + i + nl * j] = s[j + nl * i];
}

/* h1mg_extrude (src/hsmg.f:1246-1283) on extended arrays of nx^3 per
 * element: face points (tangential indices 1..nx-2) of layers l1 and
 * nx-1-l1 become f1*arr1 + f2*arr2(layer l2 / nx-1-l2), x faces, then y,
 * then z. */
static void extrude(double *a1, int l1, double f1, const double *a2, int l2, double f2, int nx,
                    int nelt)
{
    int64_t n3 = (int64_t)nx * nx * nx;
#define IX(i, j, k) ((i) + nx * (j) + (int64_t)nx * nx * (k))
    for (int e = 0; e < nelt; ++e) {
        double *u = a1 + n3 * e;
        const double *v = a2 + n3 * e;
        for (int k = 1; k <= nx - 2; ++k)
            for (int j = 1; j <= nx - 2; ++j) {
                u[IX(l1, j, k)] = f1 * u[IX(l1, j, k)] + f2 * v[IX(l2, j, k)];
                u[IX(nx - 1 - l1, j, k)] = f1 * u[IX(nx - 1 - l1, j, k)] + f2 * v[IX(nx - 1 - l2, j, k)];
            }
        for (int k = 1; k <= nx - 2; ++k)
            for (int i = 1; i <= nx - 2; ++i) {
                u[IX(i, l1, k)] = f1 * u[IX(i, l1, k)] + f2 * v[IX(i, l2, k)];
                u[IX(i, nx - 1 - l1, k)] = f1 * u[IX(i, nx - 1 - l1, k)] + f2 * v[IX(i, nx - 1 - l2, k)];
            }
        for (int j = 1; j <= nx - 2; ++j)
            for (int i = 1; i <= nx - 2; ++i) {
                u[IX(i, j, l1)] = f1 * u[IX(i, j, l1)] + f2 * v[IX(i, j, l2)];
                u[IX(i, j, nx - 1 - l1)] = f1 * u[IX(i, j, nx - 1 - l1)] + f2 * v[IX(i, j, nx - 1 - l2)];
            }
    }
#undef IX
}

/* h1mg_schwarz_toext3d / toreg3d (src/hsmg.f:1207-1244). */
static void toext3d(double *a, const double *b, int n, int nelt)
{
    int nl = n + 2;
    int64_t n3 = (int64_t)n * n * n, l3 = (int64_t)nl * nl * nl;
    memset(a, 0, sizeof *a * (size_t)(l3 * nelt));
    for (int e = 0; e < nelt; ++e)
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    a[l3 * e + (i + 1) + nl * (j + 1) + (int64_t)nl * nl * (k + 1)] =
                        b[n3 * e + i + n * j + (int64_t)n * n * k];
}

static void toreg3d(double *b, const double *a, int n, int nelt)
{
    int nl = n + 2;
    int64_t n3 = (int64_t)n * n * n, l3 = (int64_t)nl * nl * nl;
    for (int e = 0; e < nelt; ++e)
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    b[n3 * e + i + n * j + (int64_t)n * n * k] =
                        a[l3 * e + (i + 1) + nl * (j + 1) + (int64_t)nl * nl * (k + 1)];
}

/* h1mg_tnsr3d_el (src/hsmg.f:1322-1338): v = [C x B x A] u on one
 * element, A (nv,nu), Bt and Ct (nu,nv), through mxm as the reference. */
static void tnsr3d_el(double *v, int nv, const double *u, int nu, const double *A,
                      const double *Bt, const double *Ct)
{
    double work[MAXNX * MAXNX * MAXNX], work2[MAXNX * MAXNX * MAXNX];
    mxm(A, nv, u, nu, work, nu * nu);
    for (int i = 0; i < nu; ++i) mxm(work + nv * nu * i, nv, Bt, nu, work2 + nv * nv * i, nv);
    mxm(work2, nv * nv, Ct, nu, v, nv);
}

static void hsmg_free(struct orc_hsmg *M)
{
    if (!M) return;
    for (int l = 1; l <= 3; ++l) {
        mglevel *L = &M->L[l];
        free(L->gs.off);
        free(L->gs.idx);
        free(L->gss.off);
        free(L->gss.idx);
        free(L->rwt);
        free(L->fs);
        free(L->fd);
        free(L->swt);
        free(L->mask);
    }
    free(M->e);
    free(M->w);
    free(M->r);
    free(M->work);
    free(M->cainv);
    free(M->ccidx);
    free(M);
}

/* Replace the coarse step's result QQ^T r by Q Ainv Q^T r (the direct
 * coarse solve the library offers; not in the reference).  cidx gives
 * the unknown of every level-1 point (element-major, reference order),
 * -1 on the mask; ainv is the caller's dense inverse. */
int orc_hsmg_set_coarse(orc_problem *P, const double *ainv, int nc, const int32_t *cidx)
{
    struct orc_hsmg *M = P->mg;
    if (!M || nc <= 0) return 1;
    free(M->cainv);
    free(M->ccidx);
    M->nc = nc;
    M->cainv = malloc(sizeof *M->cainv * (size_t)nc * (size_t)nc);
    M->ccidx = malloc(sizeof *M->ccidx * (size_t)M->L[1].n);
    memcpy(M->cainv, ainv, sizeof *M->cainv * (size_t)nc * (size_t)nc);
    memcpy(M->ccidx, cidx, sizeof *M->ccidx * (size_t)M->L[1].n);
    return 0;
}

static void coarse_direct(const struct orc_hsmg *M, double *e)
{
    const int nc = M->nc;
    const int64_t n1 = M->L[1].n;
    double *rg = calloc((size_t)nc, sizeof *rg), *xg = calloc((size_t)nc, sizeof *xg);
    for (int64_t q = n1 - 1; q >= 0; --q) /* every copy holds the same sum */
        if (M->ccidx[q] >= 0) rg[M->ccidx[q]] = e[q];
    for (int j = 0; j < nc; ++j) {
        double s = 0.0;
        for (int i = 0; i < nc; ++i) s = s + M->cainv[(size_t)i + (size_t)nc * (size_t)j] * rg[i];
        xg[j] = s;
    }
    for (int64_t q = 0; q < n1; ++q) e[q] = M->ccidx[q] >= 0 ? xg[M->ccidx[q]] : 0.0;
    free(rg);
    free(xg);
}

int orc_hsmg_levels(const orc_problem *P) { return P->mg ? P->mg->lmax : 0; }
int orc_hsmg_nh(const orc_problem *P, int l)
{
    return (P->mg && l >= 1 && l <= P->mg->lmax) ? P->mg->L[l].nh : 0;
}

/* h1mg_setup (src/hsmg.f:3-29) and everything it calls. */
int orc_hsmg_setup(orc_problem *P)
{
    const int nx1 = P->nx1, nelt = P->nelt;
    if (nx1 < 4) return 1;
    hsmg_free(P->mg);
    struct orc_hsmg *M = calloc(1, sizeof *M);
    P->mg = M;
    /* h1mg_setup_mg_nx (:45-97).  For nx1 > 10 the reference computes
     * 2*(lx2/4)+1 from lx2, which its SIZE files do not define; any
     * lx2 >= 4 gives 3 after the min(3,.) below, taken here. */
    M->lmax = (nx1 == 4) ? 2 : 3;
    static const int mgn2[10] = {1, 2, 2, 2, 2, 3, 3, 5, 5, 5};
    int mgnx2 = 3;
    if (nx1 <= 10) mgnx2 = mgn2[nx1 - 1];
    if (nx1 == 8) mgnx2 = 3;
    if (mgnx2 > 3) mgnx2 = 3;
    int nxl[4] = {0, 1, mgnx2, mgnx2 + 1};
    nxl[M->lmax] = nx1 - 1;
    for (int l = 1; l <= M->lmax; ++l) {
        mglevel *L = &M->L[l];
        L->nx = nxl[l];
        L->nh = nxl[l] + 1;
        L->nl = L->nh + 2;
        L->n = (int64_t)L->nh * L->nh * L->nh * nelt;
        /* h1mg_setup_semhat (:99-118) */
        orc_semhat(L->ah, L->bh, L->ch, L->dh, L->zh, L->nx);
    }
    /* h1mg_setup_intp (:120-152) */
    for (int l = 1; l < M->lmax; ++l) {
        mglevel *L = &M->L[l], *F = &M->L[l + 1];
        int nf = F->nh, nc = L->nh;
        double w[2 * MAXNX + 2];
        for (int i = 0; i < nf; ++i) {
            orc_fd_weights_full(F->zh[i], L->zh, nc - 1, 1, w);
            for (int j = 0; j < nc; ++j) L->jh[i + nf * j] = w[j];
        }
        for (int i = 0; i < nf; ++i)
            for (int j = 0; j < nc; ++j) L->jht[j + nc * i] = L->jh[i + nf * j];
    }
    /* h1mg_setup_dssum (:154-169) */
    for (int l = 1; l <= M->lmax; ++l) {
        mglevel *L = &M->L[l];
        int64_t *glo = vert_box_nx(P, L->nh);
        gs_from_glo(glo, L->n, &L->gs);
        free(glo);
        glo = vert_box_nx(P, L->nl);
        gs_from_glo(glo, (int64_t)L->nl * L->nl * L->nl * nelt, &L->gss);
        free(glo);
    }
    int64_t nlmax = M->L[M->lmax].nl;
    M->work = calloc((size_t)(2 * nlmax * nlmax * nlmax * nelt), sizeof *M->work);
    M->e = calloc((size_t)(2 * P->n), sizeof *M->e);
    M->w = calloc((size_t)P->n, sizeof *M->w);
    M->r = calloc((size_t)P->n, sizeof *M->r);
    /* h1mg_setup_wtmask / h1mg_setup_rstr_wt (:171-249) */
    for (int l = 1; l <= M->lmax; ++l) {
        mglevel *L = &M->L[l];
        int n = L->nh;
        int64_t n3 = (int64_t)n * n * n;
        double *w = M->work;
        memset(w, 0, sizeof *w * (size_t)L->n);
        for (int e = 0; e < nelt; ++e)
            for (int k = 0; k < n; ++k)
                for (int j = 0; j < n; ++j)
                    for (int i = 0; i < n; ++i)
                        if (i == 0 || j == 0 || k == 0 || i == n - 1 || j == n - 1 || k == n - 1)
                            w[n3 * e + i + n * j + (int64_t)n * n * k] = 1.0;
        gs_apply(&L->gs, w);
        L->rwt = malloc(sizeof *L->rwt * (size_t)(n * n * 6 * nelt));
#define W_(i, j, k) w[n3 * e + (i) + n * (j) + (int64_t)n * n * (k)]
#define RW(a, b, s, d) L->rwt[(a) + n * ((b) + n * ((s) + 2 * ((d) + 3 * (int64_t)e)))]
        for (int e = 0; e < nelt; ++e) {
            for (int k = 0; k < n; ++k)
                for (int j = 0; j < n; ++j) {
                    RW(j, k, 0, 0) = 1.0 / W_(0, j, k);
                    RW(j, k, 1, 0) = 1.0 / W_(n - 1, j, k);
                }
            for (int k = 0; k < n; ++k)
                for (int i = 0; i < n; ++i) {
                    RW(i, k, 0, 1) = 1.0 / W_(i, 0, k);
                    RW(i, k, 1, 1) = 1.0 / W_(i, n - 1, k);
                }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    RW(i, j, 0, 2) = 1.0 / W_(i, j, 0);
                    RW(i, j, 1, 2) = 1.0 / W_(i, j, n - 1);
                }
        }
#undef RW
#undef W_
    }
    /* h1mg_setup_fdm / h1mg_setup_fast (:251-376).  Neighbour widths
     * through the fine-level dssum as the reference does. */
    double *lw = calloc((size_t)P->n, sizeof *lw);
    int64_t f3 = (int64_t)nx1 * nx1 * nx1;
#define LF(i, j, k, e) lw[(i) + nx1 * (j) + (int64_t)nx1 * nx1 * (k) + f3 * (e)]
    const double lm = 2.0;
    for (int e = 0; e < nelt; ++e) {
        LF(0, 1, 1, e) = lm;
        LF(nx1 - 1, 1, 1, e) = lm;
        LF(1, 0, 1, e) = lm;
        LF(1, nx1 - 1, 1, e) = lm;
        LF(1, 1, 0, e) = lm;
        LF(1, 1, nx1 - 1, e) = lm;
    }
    orc_dssum(P, lw);
    for (int l = 2; l <= M->lmax; ++l) {
        mglevel *L = &M->L[l];
        int nl = L->nl;
        int64_t l3 = (int64_t)nl * nl * nl;
        L->fs = malloc(sizeof *L->fs * (size_t)(nl * nl * 6 * nelt));
        L->fd = malloc(sizeof *L->fd * (size_t)(l3 * nelt));
        for (int e = 0; e < nelt; ++e) {
            double llr = LF(0, 1, 1, e) - lm, lrr = LF(nx1 - 1, 1, 1, e) - lm;
            double lls = LF(1, 0, 1, e) - lm, lrs = LF(1, nx1 - 1, 1, e) - lm;
            double llt = LF(1, 1, 0, e) - lm, lrt = LF(1, 1, nx1 - 1, e) - lm;
            int f[6];
            boundary_faces(P, e, f); /* get_bcs, src/hsmg.f:534-570 */
            double lr[MAXNX], ls[MAXNX], lt[MAXNX];
            double *s = L->fs + (int64_t)nl * nl * 6 * e;
            fast1d(s, lr, nl, f[0], f[1], llr, lm, lrr, L->ah, L->bh, L->nx);
            fast1d(s + 2 * nl * nl, ls, nl, f[2], f[3], lls, lm, lrs, L->ah, L->bh, L->nx);
            fast1d(s + 4 * nl * nl, lt, nl, f[4], f[5], llt, lm, lrt, L->ah, L->bh, L->nx);
            /* vlmax(lr(2),nr-2): largest of the eigenvalues 2..nl-1 */
            double mr = lr[1], ms = ls[1], mt = lt[1];
            for (int i = 2; i <= nl - 2; ++i) {
                if (lr[i] > mr) mr = lr[i];
                if (ls[i] > ms) ms = ls[i];
                if (lt[i] > mt) mt = lt[i];
            }
            double eps = 1.e-5 * (mr + ms + mt);
            double *d = L->fd + l3 * e;
            int64_t il = 0;
            for (int k = 0; k < nl; ++k)
                for (int j = 0; j < nl; ++j)
                    for (int i = 0; i < nl; ++i) {
                        double diag = lr[i] + ls[j] + lt[k];
                        d[il++] = (diag > eps) ? 1.0 / diag : 0.0;
                    }
        }
    }
#undef LF
    free(lw);
    /* h1mg_setup_schwarz_wt / _wt3d_1 / _wt3d_2 (:572-689) */
    for (int l = 2; l <= M->lmax; ++l) {
        mglevel *L = &M->L[l];
        int n = L->nh, nl = L->nl;
        int64_t ns = (int64_t)nl * nl * nl * nelt;
        double *w0 = M->work, *wi = M->work + ns;
        for (int64_t i = 0; i < ns; ++i) wi[i] = 1.0;
        extrude(w0, 0, 0.0, wi, 0, 1.0, nl, nelt);
        gs_apply(&L->gss, wi);
        extrude(wi, 0, 1.0, w0, 0, -1.0, nl, nelt);
        extrude(wi, 2, 1.0, wi, 0, 1.0, nl, nelt);
        toreg3d(w0, wi, n, nelt);
        gs_apply(&L->gs, w0);
        int64_t n3 = (int64_t)n * n * n;
        L->swt = malloc(sizeof *L->swt * (size_t)(n * n * 12 * nelt));
#define W_(i, j, k) w0[n3 * e + (i) + n * (j) + (int64_t)n * n * (k)]
#define SW(a, b, s, d) L->swt[(a) + n * ((b) + n * ((s) + 4 * ((d) + 3 * (int64_t)e)))]
        for (int e = 0; e < nelt; ++e) {
            for (int k = 0; k < n; ++k)
                for (int j = 0; j < n; ++j) {
                    SW(j, k, 0, 0) = 1.0 / W_(0, j, k);
                    SW(j, k, 1, 0) = 1.0 / W_(1, j, k);
                    SW(j, k, 2, 0) = 1.0 / W_(n - 2, j, k);
                    SW(j, k, 3, 0) = 1.0 / W_(n - 1, j, k);
                }
            for (int k = 0; k < n; ++k)
                for (int i = 0; i < n; ++i) {
                    SW(i, k, 0, 1) = 1.0 / W_(i, 0, k);
                    SW(i, k, 1, 1) = 1.0 / W_(i, 1, k);
                    SW(i, k, 2, 1) = 1.0 / W_(i, n - 2, k);
                    SW(i, k, 3, 1) = 1.0 / W_(i, n - 1, k);
                }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    SW(i, j, 0, 2) = 1.0 / W_(i, j, 0);
                    SW(i, j, 1, 2) = 1.0 / W_(i, j, 1);
                    SW(i, j, 2, 2) = 1.0 / W_(i, j, n - 2);
                    SW(i, j, 3, 2) = 1.0 / W_(i, j, n - 1);
                }
        }
#undef SW
#undef W_
    }
    /* mg_set_msk / h1mg_setup_mask (:824-911): every domain-boundary
     * point of each level, element by element. */
    for (int l = 1; l <= M->lmax; ++l) {
        mglevel *L = &M->L[l];
        int n = L->nh;
        int64_t n3 = (int64_t)n * n * n, cnt = 0;
        L->mask = malloc(sizeof *L->mask * (size_t)L->n);
        for (int e = 0; e < nelt; ++e) {
            int f[6];
            boundary_faces(P, e, f);
            for (int k = 0; k < n; ++k)
                for (int j = 0; j < n; ++j)
                    for (int i = 0; i < n; ++i)
                        if (on_boundary(f, i, j, k, n))
                            L->mask[cnt++] = (int32_t)(n3 * e + i + n * j + (int64_t)n * n * k);
        }
        L->nmask = cnt;
    }
    return 0;
}

static void mg_mask(const mglevel *L, double *u)
{
    for (int64_t q = 0; q < L->nmask; ++q) u[L->mask[q]] = 0.;
}

/* h1mg_schwarz_wt3d2 (src/hsmg.f:1144-1178): e *= sqrt(wt) on the two
 * outer layers of each direction, each point once. */
static void schwarz_wt(const mglevel *L, double *e, int nelt)
{
    int n = L->nh;
    int64_t n3 = (int64_t)n * n * n;
#define E_(i, j, k) e[n3 * ie + (i) + n * (j) + (int64_t)n * n * (k)]
#define SW(a, b, s, d) L->swt[(a) + n * ((b) + n * ((s) + 4 * ((d) + 3 * (int64_t)ie)))]
    for (int ie = 0; ie < nelt; ++ie) {
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j) {
                E_(0, j, k) = E_(0, j, k) * sqrt(SW(j, k, 0, 0));
                E_(1, j, k) = E_(1, j, k) * sqrt(SW(j, k, 1, 0));
                E_(n - 2, j, k) = E_(n - 2, j, k) * sqrt(SW(j, k, 2, 0));
                E_(n - 1, j, k) = E_(n - 1, j, k) * sqrt(SW(j, k, 3, 0));
            }
        for (int k = 0; k < n; ++k)
            for (int i = 2; i <= n - 3; ++i) {
                E_(i, 0, k) = E_(i, 0, k) * sqrt(SW(i, k, 0, 1));
                E_(i, 1, k) = E_(i, 1, k) * sqrt(SW(i, k, 1, 1));
                E_(i, n - 2, k) = E_(i, n - 2, k) * sqrt(SW(i, k, 2, 1));
                E_(i, n - 1, k) = E_(i, n - 1, k) * sqrt(SW(i, k, 3, 1));
            }
        for (int j = 2; j <= n - 3; ++j)
            for (int i = 2; i <= n - 3; ++i) {
                E_(i, j, 0) = E_(i, j, 0) * sqrt(SW(i, j, 0, 2));
                E_(i, j, 1) = E_(i, j, 1) * sqrt(SW(i, j, 1, 2));
                E_(i, j, n - 2) = E_(i, j, n - 2) * sqrt(SW(i, j, 2, 2));
                E_(i, j, n - 1) = E_(i, j, n - 1) * sqrt(SW(i, j, 3, 2));
            }
    }
#undef SW
#undef E_
}

/* h1mg_schwarz (src/hsmg.f:1036-1050) with h1mg_schwarz_part1
 * (:1052-1095) and h1mg_fdm / h1mg_do_fast (:1286-1320); sigma = 1.
 * The reference also weights e before part1, which part1 then
 * overwrites entirely; that step has no effect and is omitted. */
static void mg_schwarz(struct orc_hsmg *M, const mglevel *L, double *e, double *r, int nelt)
{
    int n = L->nh, nl = L->nl;
    int64_t l3 = (int64_t)nl * nl * nl, ns = l3 * nelt;
    double *w0 = M->work, *w1 = M->work + ns;
    mg_mask(L, r);
    toext3d(w0, r, n, nelt);
    extrude(w0, 0, 0.0, w0, 2, 1.0, nl, nelt);
    gs_apply(&L->gss, w0);
    extrude(w0, 0, 1.0, w0, 2, -1.0, nl, nelt);
    for (int ie = 0; ie < nelt; ++ie) {
        const double *s = L->fs + (int64_t)nl * nl * 6 * ie;
        const double *Sx = s, *SxT = s + nl * nl, *Sy = s + 2 * nl * nl, *SyT = s + 3 * nl * nl,
                     *Sz = s + 4 * nl * nl, *SzT = s + 5 * nl * nl;
        double *ee = w1 + l3 * ie, *rr = w0 + l3 * ie;
        const double *d = L->fd + l3 * ie;
        tnsr3d_el(ee, nl, rr, nl, SxT, Sy, Sz);
        for (int64_t i = 0; i < l3; ++i) rr[i] = d[i] * ee[i];
        tnsr3d_el(ee, nl, rr, nl, Sx, SyT, SzT);
    }
    extrude(w0, 0, 0.0, w1, 0, 1.0, nl, nelt);
    gs_apply(&L->gss, w1);
    extrude(w1, 0, 1.0, w0, 0, -1.0, nl, nelt);
    extrude(w1, 2, 1.0, w1, 0, 1.0, nl, nelt);
    toreg3d(e, w1, n, nelt);
    gs_apply(&L->gs, e);
    mg_mask(L, e);
    schwarz_wt(L, e, nelt);
    /* cmult(e, sigma=1) leaves e unchanged */
}

/* h1mg_rstr (src/hsmg.f:1340-1357): r := J^T (wt .* r) in place, from
 * level l+1 to level l, then optionally dssum on level l. */
static void mg_rstr(const mglevel *C, const mglevel *F, double *r, int ifdssum, int nelt)
{
    int n = F->nh;
    int64_t n3 = (int64_t)n * n * n;
#define U_(i, j, k) r[n3 * ie + (i) + n * (j) + (int64_t)n * n * (k)]
#define RW(a, b, s, d) F->rwt[(a) + n * ((b) + n * ((s) + 2 * ((d) + 3 * (int64_t)ie)))]
    for (int ie = 0; ie < nelt; ++ie) { /* h1mg_do_wt (:1359-1388) */
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j) {
                U_(0, j, k) = U_(0, j, k) * RW(j, k, 0, 0);
                U_(n - 1, j, k) = U_(n - 1, j, k) * RW(j, k, 1, 0);
            }
        for (int k = 0; k < n; ++k)
            for (int i = 1; i <= n - 2; ++i) {
                U_(i, 0, k) = U_(i, 0, k) * RW(i, k, 0, 1);
                U_(i, n - 1, k) = U_(i, n - 1, k) * RW(i, k, 1, 1);
            }
        for (int j = 1; j <= n - 2; ++j)
            for (int i = 1; i <= n - 2; ++i) {
                U_(i, j, 0) = U_(i, j, 0) * RW(i, j, 0, 2);
                U_(i, j, n - 1) = U_(i, j, n - 1) * RW(i, j, 1, 2);
            }
    }
#undef RW
#undef U_
    /* h1mg_tnsr1_3d (:1402-1434) in place, elements in increasing order
     * since the coarse result is smaller */
    int nv = C->nh, nu = F->nh;
    int64_t nu3 = (int64_t)nu * nu * nu, nv3 = (int64_t)nv * nv * nv;
    double tmp[MAXNX * MAXNX * MAXNX];
    for (int ie = 0; ie < nelt; ++ie) {
        tnsr3d_el(tmp, nv, r + nu3 * ie, nu, C->jht, C->jh, C->jh);
        memcpy(r + nv3 * ie, tmp, sizeof *tmp * (size_t)nv3);
    }
    if (ifdssum) gs_apply(&C->gs, r);
}

/* h1mg_intp (src/hsmg.f:1436-1443): uf = [J x J x J] uc */
static void mg_intp(const mglevel *C, const mglevel *F, double *uf, const double *uc, int nelt)
{
    int nv = F->nh, nu = C->nh;
    int64_t nu3 = (int64_t)nu * nu * nu, nv3 = (int64_t)nv * nv * nv;
    for (int ie = 0; ie < nelt; ++ie)
        tnsr3d_el(uf + nv3 * ie, nv, uc + nu3 * ie, nu, C->jh, C->jht, C->jht);
}

/* h1mg_solve (src/hsmg.f:960-1034) */
void orc_hsmg_solve(orc_problem *P, double *z, double *rhs)
{
    struct orc_hsmg *M = P->mg;
    const int nelt = P->nelt, lmax = M->lmax;
    double *e = M->e, *w = M->w, *r = M->r;
    mglevel *L = M->L;
    mg_schwarz(M, &L[lmax], z, rhs, nelt);
    int64_t n = L[lmax].n, is = 0;
    memcpy(r, rhs, sizeof *r * (size_t)n);
    for (int l = lmax - 1; l >= 2; --l) {
        is += n;
        n = L[l].n;
        mg_rstr(&L[l], &L[l + 1], r, 1, nelt);
        mg_schwarz(M, &L[l], e + is, r, nelt);
    }
    is += n;
    mg_rstr(&L[1], &L[2], r, 0, nelt);
    mg_mask(&L[1], r);
    memcpy(e + is, r, sizeof *e * (size_t)L[1].n); /* hsmg_coarse_solve (:945-957) */
    gs_apply(&L[1].gs, e + is);
    if (M->cainv) coarse_direct(M, e + is);
    mg_mask(&L[1], e + is);
    for (int l = 2; l <= lmax - 1; ++l) {
        int64_t im = is;
        is -= n;
        n = L[l].n;
        mg_intp(&L[l - 1], &L[l], w, e + im, nelt);
        for (int64_t i = 0; i < n; ++i) e[is + i] = e[is + i] + w[i];
    }
    n = L[lmax].n;
    mg_intp(&L[lmax - 1], &L[lmax], w, e + is, nelt);
    for (int64_t i = 0; i < n; ++i) z[i] = z[i] + w[i];
}
