/* Oracle: a CPU restatement of Nekbone's CG Poisson hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or the timed CPU baseline).  The product path never calls it.
 *
 * Every function states the reference file:line it restates; evaluation
 * order (left-to-right sums, product grouping) follows the reference so
 * that results agree bit for bit with the flang-compiled reference, which
 * the tests check through fixtures under tests/golden/.
 *
 * Layout: all point arrays are the reference's u(i,j,k,e), i fastest,
 * 0-based local index l = i + nx*j + nx*nx*k + nx^3*e.  Geometric factors
 * are the reference's AoS g(6, l).
 */
#ifndef NEKBONE_ORACLE_H
#define NEKBONE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem orc_problem;

enum { ORC_PRECOND_IDENTITY = 0, ORC_PRECOND_JACOBI = 1, ORC_PRECOND_HSMG = 2 };
enum { ORC_MASK_REF = 0, ORC_MASK_DIRICHLET = 1 };

/* GLL points/weights (src/speclib.f:107-518) and 1-D operators
 * (src/semhat.f:2-93). a,d are (n+1)x(n+1) column-major. */
void orc_zwgll(double *z, double *w, int np);
void orc_fd_weights_full(double xx, const double *x, int n, int m, double *c);
void orc_semhat(double *a, double *b, double *c, double *d, double *z, int n);

/* Box factorisation (src/driver.f:196-247). */
void orc_cubic(int *mx, int *my, int *mz, int np);

/* Problem set-up as src/driver.f:117-194 (np=1 box), prox_dssum.f:14-73,
 * driver.f:274-291, prox_setup.f:2-52. */
orc_problem *orc_create(int nelx, int nely, int nelz, int nx1);
void orc_free(orc_problem *P);
int64_t orc_npts(const orc_problem *P);
void orc_get_dxm1(const orc_problem *P, double *d);          /* nx1^2 */
void orc_get_c(const orc_problem *P, double *c);             /* n     */
void orc_get_g(const orc_problem *P, double *g);             /* 6n    */
void orc_set_g(orc_problem *P, const double *g);             /* 6n    */
void orc_get_glo(const orc_problem *P, int64_t *glo);        /* n     */
int64_t orc_gs_ngroups(const orc_problem *P);
int64_t orc_gs_nshared(const orc_problem *P);
/* Group offsets (ngroups+1) and member indices (nshared), jl order. */
void orc_gs_map(const orc_problem *P, int64_t *off, int32_t *idx);

/* Mask lists (src/cg.f:214-368): full Dirichlet boundary list and the
 * list the reference records on its first maskit call (element 1's
 * boundary positions).  Returns list lengths. */
int64_t orc_mask_full_count(const orc_problem *P);
int64_t orc_mask_ref_count(const orc_problem *P);
void orc_mask_lists(const orc_problem *P, int32_t *full, int32_t *ref);
void orc_set_mask_mode(orc_problem *P, int mode);
void orc_reset_mask(orc_problem *P); /* cmask(-1)=1, src/driver.f:125 */

/* Operators. */
void orc_dssum(const orc_problem *P, double *u);             /* prox_dssum.f:2-12 */
void orc_ax_e(const orc_problem *P, int e, double *w, const double *u);
void orc_ax(orc_problem *P, double *w, const double *u);     /* cg.f:96-125 */
void orc_maskit(orc_problem *P, double *w);                  /* cg.f:214-258 */
double orc_glsc3(const double *a, const double *b, const double *m, int64_t n);
/* the same terms, summed with compensation (correct to an ulp) */
double orc_glsc3_exact(const double *a, const double *b, const double *m, int64_t n);
/* 1: orc_cg sums its dot products with orc_glsc3_exact */
void orc_set_exact_dots(orc_problem *P, int on);

/* Jacobi preconditioner (not in the reference): d = QQ^T diag(A_L) + 0.1,
 * dinv = 1/d.  diag(A_L) evaluated in the fixed order documented in the
 * source. */
void orc_jacobi_diag_local(const orc_problem *P, double *dl);
void orc_setup_jacobi(orc_problem *P);
void orc_get_dinv(const orc_problem *P, double *dinv);

/* CG (src/cg.f:2-84).  hist receives every glsc3 value in call order:
 * rnorm^2, then (rtz1, pap, rtr) per iteration: 1+3*niter doubles.
 * x, r are n-vectors; f is the masked-on-entry right-hand side. */
void orc_cg(orc_problem *P, double *x, const double *f, double *r,
            int niter, int precond, double *hist);

/* HSMG preconditioner (src/hsmg.f): h1mg_setup (:3-29) on the problem's
 * current mesh, and one application z = M^-1 rhs, h1mg_solve (:960-1034),
 * which also zeroes the Dirichlet points of rhs in place as the reference
 * does (:1067).  Needs nx1 >= 4.  Returns 0 on success. */
int orc_hsmg_setup(orc_problem *P);
void orc_hsmg_solve(orc_problem *P, double *z, double *rhs);
/* Level data for checks: number of levels, and points per direction of
 * level l (1 = coarsest). */
int orc_hsmg_levels(const orc_problem *P);
int orc_hsmg_nh(const orc_problem *P, int l);
/* Coarse step with a caller-given dense inverse (checks the library's
 * optional direct coarse solve; no counterpart in the reference):
 * e = Q ainv Q^T r on level 1, cidx = unknown of each level-1 point
 * (element-major, reference order), -1 on the mask.  Returns 0 on
 * success. */
int orc_hsmg_set_coarse(orc_problem *P, const double *ainv, int nc, const int32_t *cidx);

/* Synthetic inputs. */
void orc_rand_uniform(double *f, int64_t n, uint64_t seed); /* U[-1,1) */
void orc_rhs_from_raw(const orc_problem *P, double *f);     /* dssum, then *c */

/* Deformed geometry (not in the reference): vertices of the nelx x nely x
 * nelz box of width-2 elements, each coordinate perturbed by
 * amp*2*U[-1,1) (seeded), trilinear map, g = w_i w_j w_k detJ J^-1 J^-T. */
void orc_deformed_g(orc_problem *P, double amp, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
