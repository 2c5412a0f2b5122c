/* nekbone_amd: MI355X (gfx950) implementation of Nekbone's CG Poisson hot
 * path behind a C ABI.
 *
 * Two layers:
 *
 *  1. Reference-named, Fortran-callable entry points (trailing underscore,
 *     every argument by reference, REAL = double, INTEGER = int32,
 *     glo_num = int64).  A Fortran host links these in place of the
 *     reference's objects; they act on the problem bound by
 *     nekbone_setup_() (the state the reference keeps in COMMON blocks).
 *
 *  2. Explicit nekbone_* entry points with an opaque problem handle, plain
 *     pointers and sizes, returning an int status (0 = success; on failure
 *     nekbone_last_error() describes it).  *_d variants take device
 *     pointers; the others take host arrays and copy.
 *
 * Citations are to the reference sources (src/...).
 */
#ifndef NEKBONE_AMD_H
#define NEKBONE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Status, lifecycle                                                        */
/* ---------------------------------------------------------------------- */
enum {
    NEKBONE_OK = 0,
    NEKBONE_ERR_ARG = 1,      /* bad argument / size                    */
    NEKBONE_ERR_HIP = 2,      /* HIP runtime error                      */
    NEKBONE_ERR_STATE = 3,    /* call order (e.g. no Jacobi set up)     */
    NEKBONE_ERR_NUMERIC = 4,  /* non-finite or non-positive CG scalar   */
    NEKBONE_ERR_UNSUPPORTED = 5,
    /* a tolerance solve stopped at maxiter with rtr > tol^2 rtr_0; its
     * outputs (x, r, iteration count, final rtr) are written all the same */
    NEKBONE_NOT_CONVERGED = 6
};

enum { NEKBONE_PRECOND_IDENTITY = 0, NEKBONE_PRECOND_JACOBI = 1, NEKBONE_PRECOND_HSMG = 2 };
enum { NEKBONE_MASK_REF = 0, NEKBONE_MASK_DIRICHLET = 1 };
/* Arithmetic of the element operator and vector updates:
 *   EXACT  FP64, the reference's evaluation order, no FMA contraction
 *          (bit-identical ax/dssum; CG differs only through glsc3's
 *          summation order).
 *   FMA    FP64 with fused multiply-adds (same order of terms).           */
enum { NEKBONE_ARITH_EXACT = 0, NEKBONE_ARITH_FMA = 1 };

/* Last error message of the calling thread ("" if none). */
const char *nekbone_last_error(void);
/* Select the device (default 0) and create the library's stream. */
int nekbone_init(int device);
int nekbone_finalize(void);

typedef struct nekbone_problem nekbone_problem;

/* ---------------------------------------------------------------------- */
/* Problem set-up                                                           */
/* ---------------------------------------------------------------------- */

/* Box of nelx*nely*nelz elements with nx1 GLL points per direction, set up
 * as the reference driver does for one rank: element numbering
 * (src/driver.f:117-194, lglel(e)=e), global numbering and gather-scatter
 * map (src/prox_dssum.f:14-73, src/jl/gs.c:88-113,330-357), inverse
 * multiplicity c (src/driver.f:274-291), D = dxm1 (src/prox_setup.f:2-25,
 * src/semhat.f), brick geometric factors (src/prox_setup.f:27-52), mask
 * lists (src/cg.f:214-368).  2 <= nx1 <= 16. */
int nekbone_problem_create_box(nekbone_problem **P, int nelx, int nely, int nelz, int nx1);

/* Drop-in set-up from the caller's arrays: the host has already run the
 * reference's set-up.  dxm1(nx1,nx1) column-major; glo_num(n) int64 as
 * set_vert_box produces; gxyz(6,n) AoS; c(n); lglel(nelt) 1-based global
 * element numbers of the nelx x nely x nelz box (for the Dirichlet faces,
 * src/cg.f:302-368). */
int nekbone_problem_create(nekbone_problem **P, int nelt, int nx1, int nelx, int nely, int nelz,
                           const int32_t *lglel, const double *dxm1, const int64_t *glo_num,
                           const double *gxyz, const double *c);
int nekbone_problem_destroy(nekbone_problem *P);

int64_t nekbone_problem_npts(const nekbone_problem *P);
int nekbone_problem_nx1(const nekbone_problem *P);
int nekbone_problem_nelt(const nekbone_problem *P);

/* Replace the geometric factors: from host AoS g(6,n), or by the deformed
 * trilinear box (vertices perturbed by amp*2*U[-1,1), seeded; computed on
 * the device).  The deformed map is not in the reference; it reduces
 * exactly to src/prox_setup.f:27-52 for amp = 0. */
int nekbone_set_geometry(nekbone_problem *P, const double *gxyz);
int nekbone_set_geometry_deformed(nekbone_problem *P, double amp, uint64_t seed);
int nekbone_get_geometry(nekbone_problem *P, double *gxyz);

/* Host copies of set-up data (for checks). */
int nekbone_get_dxm1(const nekbone_problem *P, double *dxm1);
int nekbone_get_c(nekbone_problem *P, double *c);
int nekbone_get_glo(const nekbone_problem *P, int64_t *glo);
int64_t nekbone_gs_ngroups(const nekbone_problem *P);
int64_t nekbone_gs_nshared(const nekbone_problem *P);

/* Mask: REF reproduces src/cg.f:214-258 exactly (the first maskit after a
 * reset zeroes every domain-boundary point and records element 1's
 * positions; later calls zero only those).  DIRICHLET zeroes every
 * domain-boundary point on every call.  reset = cmask(-1)=1
 * (src/driver.f:125). */
int nekbone_set_mask_mode(nekbone_problem *P, int mode);
int nekbone_reset_mask(nekbone_problem *P);
int nekbone_set_arith(nekbone_problem *P, int arith);
/* CG kernel sequence: 1 (default when c = 1/multiplicity) fuses the
 * direction and x updates, the Helmholtz shift, the mask and pap into the
 * element and gather-scatter kernels; 0 runs the reference's sequence of
 * separate passes.  Values agree bit for bit between the two; only the
 * dot products' summation trees differ. */
int nekbone_set_fused(nekbone_problem *P, int on);
/* Form of the fused CG iteration.  on = 1 (default): where the problem's
 * shared-point map has the entity form of a hex mesh (every group's
 * members at the same offset within matching vertices, edges or faces)
 * and the full Dirichlet mask applies, the vector-update kernel sums the
 * copies of each shared point itself (in the reference's member order)
 * and the element kernel forms pap from its own copies, so an iteration
 * runs two vector kernels instead of three; on = 0: the separate
 * gather-scatter kernel; on = -1: leave the setting.  *active (if not
 * null) receives 1 when the gather form applies to the next solve.  Both
 * forms compute w with the same operations in the same order; only the
 * pap sum (hence alpha) is accumulated in another order.  No counterpart
 * in the reference, whose dssum is a separate pass (src/cg.f:116). */
int nekbone_set_gather(nekbone_problem *P, int on, int *active);
/* x update of the gather form.  on = 1 (default): the directions of the
 * last 10 iterations are kept in a ring of 10 vectors (allocated on the
 * first such solve when the device has room for them, else this setting
 * has no effect) and x receives their alpha p once every 10 iterations, in
 * iteration order and with the same operations, so x is bit-identical to
 * the per-iteration update while the element kernel neither reads nor
 * writes x; on = 0: x is updated every iteration (and the ring freed).
 * Applies to the FP64 CG (the mixed solver's FP32 inner iterations update
 * x every iteration).  No counterpart in the reference (src/cg.f:63
 * updates x each iteration). */
int nekbone_set_xring(nekbone_problem *P, int on);
/* The forms the last FP64 solve of P ran in: *gather = 1 for the gather
 * form, *xring = 1 when it used the x-ring (either pointer may be null). */
int nekbone_get_cg_form(nekbone_problem *P, int *gather, int *xring);
/* 1 (default): CG iterations are replayed in chunks of 10 from a HIP graph
 * captured on first use (re-captured when buffers or settings change; not
 * while per-kernel timing is enabled); 0: every kernel launched directly.
 * Same kernels, same results. */
int nekbone_set_graph(nekbone_problem *P, int on);

/* Element-kernel shape: elements per workgroup; pf = 0 geometric factors
 * loaded where they are used, 1 prefetched at kernel entry, 2 as 0 but
 * compiled for 5 waves per SIMD, 3 loaded one t-step of the third phase
 * ahead of their use; x update carried by the element kernel
 * (0/1).  nekbone_autotune times every shape compiled for this order on
 * this problem (a few fused CG iterations each), keeps the fastest and
 * remembers it for later problems of the same order, preconditioner and
 * arithmetic in this process; *best_ms = its time per iteration. */
int nekbone_autotune(nekbone_problem *P, int precond, double *best_ms);
int nekbone_get_variant(const nekbone_problem *P, int *epb, int *pf, int *xu);
/* The autotuner's choices on disk (JSON, "nekbone_amd tuning 1"): save
 * writes every shape found in this process; load takes them back (shapes
 * not compiled for their order are skipped), so problems created later
 * start with them without tuning again.  *loaded may be NULL. */
int nekbone_tuning_save(const char *path);
int nekbone_tuning_load(const char *path, int *loaded);
int nekbone_set_variant(nekbone_problem *P, int epb, int pf, int xu);

/* Jacobi preconditioner (the reference has none): dinv = 1/(QQ^T
 * diag(A_L) + 0.1), computed on the device. */
int nekbone_setup_jacobi(nekbone_problem *P);
int nekbone_get_dinv(nekbone_problem *P, double *dinv);

/* HSMG preconditioner (src/hsmg.f): h1mg_setup (:3-29) on this problem's
 * box, then NEKBONE_PRECOND_HSMG is available to the CG calls.  Levels of
 * order 1, min(3, ...) and nx1-1 as h1mg_setup_mg_nx (:45-97) chooses;
 * needs nx1 >= 4.  nekbone_hsmg_solve applies it once, h1mg_solve
 * (:960-1034): z = M^-1 r on host arrays, and r comes back with its
 * Dirichlet points zeroed, as the reference leaves it (:1067). */
int nekbone_setup_hsmg(nekbone_problem *P);
int nekbone_hsmg_solve(nekbone_problem *P, double *z, double *r);
/* on = 1: the local FDM solves of the Schwarz smoother (levels nx1-1 and
 * the middle level) run in FP32 with FP32 copies of the eigenbases built
 * by nekbone_setup_hsmg; restriction, interpolation, the coarse level and
 * the outer CG stay FP64.  on = 0 (default): everything FP64.  The
 * reference has no FP32 path (it is this library's option). */
int nekbone_set_hsmg_fp32(nekbone_problem *P, int on);
/* HSMG coarse level (no counterpart in the reference, whose coarse step,
 * hsmg_coarse_solve src/hsmg.f:945-957, only sums the restricted
 * residual).  mode 0 (default): the reference's coarse step.  mode 1: an
 * FP64 direct solve with the Galerkin projection of the operator onto the
 * trilinear functions of the mesh vertices (A_c = sum_e J^T A_L J + 0.1
 * sum_e J^T C_e J on the unmasked vertices), by a dense inverse formed
 * once with rocSOLVER; at most 8192 coarse unknowns.  Requires
 * nekbone_setup_hsmg; a new nekbone_setup_hsmg resets the mode to 0. */
int nekbone_set_hsmg_coarse(nekbone_problem *P, int mode);

/* ---------------------------------------------------------------------- */
/* Operators (host arrays; *_d: device pointers on the library stream)     */
/* ---------------------------------------------------------------------- */
/* Every vector argument, host or device, holds the reference's point order
 * u(i,j,k,e).  The host-array calls return with their results written; the
 * *_d operator calls return with their work queued on the library stream
 * (nekbone_sync before the host reads the result by other means). */
/* w = mask(QQ^T A_L u + 0.1 u): ax, src/cg.f:96-125 */
int nekbone_ax(nekbone_problem *P, double *w, const double *u);
int nekbone_ax_d(nekbone_problem *P, double *w, const double *u);
/* w = A_L u element by element, no dssum/shift/mask: ax_e over all
 * elements, src/cg.f:142-168 */
int nekbone_ax_local(nekbone_problem *P, double *w, const double *u);
/* u = QQ^T u: dssum, src/prox_dssum.f:2-12 */
int nekbone_dssum(nekbone_problem *P, double *u);
int nekbone_dssum_d(nekbone_problem *P, double *u);
/* maskit on a host vector (advances the mask state like the reference) */
int nekbone_maskit(nekbone_problem *P, double *w);
/* sum a*b*m: glsc3, src/math.f:603-617 */
int nekbone_glsc3(nekbone_problem *P, const double *a, const double *b, const double *m,
                  double *result);

/* ---------------------------------------------------------------------- */
/* CG (src/cg.f:2-84)                                                       */
/* ---------------------------------------------------------------------- */
/* x = 0, r = mask(f), then niter PCG iterations.  hist (may be NULL)
 * receives every glsc3 value in call order: rnorm^2, then (rtz1, pap, rtr)
 * per iteration (1+3*niter doubles).  Host arrays; x, r may be NULL. */
int nekbone_cg(nekbone_problem *P, double *x, const double *f, double *r, int niter,
               int precond, double *hist);
/* Same on device arrays already resident (x, f, r device pointers of n
 * doubles; hist host, may be NULL). */
int nekbone_cg_d(nekbone_problem *P, double *x, const double *f, double *r, int niter,
                 int precond, double *hist);
/* Run to a relative residual: stops after the first iteration with
 * rtr <= tol^2 * rtr_0 (rtr_0 the initial glsc3(r,c,r)), or at maxiter.
 * *iters receives the iteration count.  Returns NEKBONE_NOT_CONVERGED
 * when it stopped at maxiter without meeting tol. */
int nekbone_cg_tol_d(nekbone_problem *P, double *x, const double *f, double *r, double tol,
                     int maxiter, int precond, int *iters, double *final_rtr);

/* Mixed precision.  No counterpart in the reference (its CG,
 * src/cg.f:2-84, is FP64 throughout); the stopping rule is that of
 * nekbone_cg_tol_d, evaluated on the FP64 residual of the FP64 operator:
 * glsc3(r,c,r) <= tol^2 glsc3(r0,c,r0) with r = mask(f - A x).
 * Inner iterations run in FP32 (geometric factors, D, vectors and
 * gather-scatter in FP32; dot products accumulated in FP64); x is
 * accumulated in FP64 and the residual is recomputed in FP64 ("reliable
 * update", the search direction is kept) whenever the FP32 residual norm
 * has dropped by the factor delta since the last update (0 selects 1e-3,
 * 5e-3 with FP16 factors, 3e-2 with the FP16 matrix-core operator; see
 * nekbone_set_mixed_inner)
 * or the FP32 iteration reports convergence.  x, f, r as for
 * nekbone_cg_tol_d.  *iters: inner iterations; *updates: FP64 residual
 * recomputations; *final_rtr: glsc3(r,c,r) of the returned x.  Returns
 * NEKBONE_NOT_CONVERGED when maxiter inner iterations ran without the FP64
 * residual meeting tol (outputs written, as for nekbone_cg_tol_d).  Requires
 * c = 1/multiplicity.  The FP32 copies of g, D and dinv are made on first
 * use and remade after the geometry or the Jacobi diagonal changes. */
int nekbone_cg_mixed_d(nekbone_problem *P, double *x, const double *f, double *r, double tol,
                       int maxiter, int precond, double delta, int *iters, int *updates,
                       double *final_rtr);
/* 1: the mixed solver's inner iterations read the geometric factors in
 * FP16 (per element and factor scaled by its largest magnitude, 12 bytes
 * per point instead of 24); the FP64 residual updates still use the FP64
 * factors.  0 (default): FP32 factors. */
int nekbone_set_mixed_g16(nekbone_problem *P, int on);
/* Inner operator of the mixed solver (no counterpart in the reference):
 * NEKBONE_MIXED_F32 (default) FP32 factors and FP32 contractions;
 * NEKBONE_MIXED_G16 as nekbone_set_mixed_g16(P, 1);
 * NEKBONE_MIXED_F16_MFMA FP16 factors and FP16 matrix-core contractions:
 * per element, the input is scaled by its largest magnitude and rounded to
 * FP16, the three gradient contractions run on MFMA (16x16x16, FP32
 * accumulation), the geometric product in FP32, the products scaled by
 * their largest magnitude and rounded to FP16 for the three transposed
 * contractions on MFMA, w accumulated in FP32.  nx1 = 4..16.  The FP64
 * residual updates and the stopping rule are unchanged. */
#define NEKBONE_MIXED_F32 0
#define NEKBONE_MIXED_G16 1
#define NEKBONE_MIXED_F16_MFMA 2
int nekbone_set_mixed_inner(nekbone_problem *P, int mode);
/* How the mixed solver's FP64 residual refresh folds the FP32 increment
 * into x (no counterpart in the reference): 1 (default) inside the FP64
 * element kernel that applies A_L to the new x; 0 in a pass of its own
 * before it (also what runs with NEKBONE_ARITH_FMA).  The arithmetic is the
 * same, so the solves agree bit for bit; 1 saves one pass over x, x32 and
 * p32 per refresh. */
int nekbone_set_mixed_refresh(nekbone_problem *P, int fused);
/* FP32 element-kernel shape of the mixed solver (as nekbone_get_variant /
 * nekbone_set_variant), and its autotuner: times every compiled shape over
 * a few inner iterations and keeps the fastest for this problem. */
int nekbone_get_variant_f32(nekbone_problem *P, int *epb, int *pf, int *xu);
int nekbone_set_variant_f32(nekbone_problem *P, int epb, int pf, int xu);
int nekbone_autotune_mixed(nekbone_problem *P, int precond, double *best_ms);
/* The FP32 element operator of the mixed solver, w = A_L u element by
 * element (host float arrays, reference point order, n each). */
int nekbone_ax_local_f32(nekbone_problem *P, float *w, const float *u);

/* Batched solves (no counterpart in the reference, whose driver runs one
 * mesh at a time, src/driver.f:28-79).  A batch holds any number of
 * problems of any sizes and orders; each keeps its own mask mode,
 * geometry and Jacobi set-up, and needs c = 1/multiplicity.  Every
 * iteration of the whole batch is one launch per kernel per order group
 * (element operator, gather-scatter) and one per kernel for all problems
 * (scalars, vector update), replayed from a captured HIP graph in chunks;
 * each problem's values are those of its own nekbone_cg_tol_d solve with
 * the batch's element-kernel shape.  The batch sets that shape on its
 * problems.  The problems must outlive the batch; each solve re-reads
 * their device buffers into the batch's tables, which the captured graphs
 * read, so set-up calls on a problem between solves are seen. */
typedef struct nekbone_batch nekbone_batch;
int nekbone_batch_create(nekbone_batch **B, nekbone_problem *const *probs, int nprob);
int nekbone_batch_destroy(nekbone_batch *B);
int nekbone_batch_size(const nekbone_batch *B);
/* 1: the deferred x update rides in the element kernel (default); 0: in
 * the vector update.  nekbone_batch_autotune times both and keeps the
 * faster; *best_ms is its time per batched iteration. */
int nekbone_batch_set_xu(nekbone_batch *B, int xu);
/* 1 (default): chunks of ten iterations are captured once into a HIP graph
 * and replayed; 0: every kernel is launched directly (for tracing tools). */
int nekbone_batch_set_graph(nekbone_batch *B, int on);
int nekbone_batch_autotune(nekbone_batch *B, int precond, double *best_ms);
/* Element-kernel path: tileset = -1 launches the element kernel once per
 * order group; 0 .. ntilesets-1 runs the problems of orders 6, 8, 10, 12
 * (N = 5, 7, 9, 11) in one persistent launch of the work-queue kernel over
 * (problem, element tile) items with that compiled tile set (elements per
 * workgroup per order), the other orders per group.  The default is -1
 * (measured faster on the 64-problem workload); nekbone_batch_autotune
 * picks the fastest path and shapes.  Every problem's values are those of its single
 * solve with the same shape (the batch sets it on the problem). */
int nekbone_batch_set_queue(nekbone_batch *B, int tileset);
int nekbone_batch_get_path(const nekbone_batch *B, int *tileset, int *xu, int *ntilesets);
/* Whether the last batched solve ran in the gather form (every member
 * eligible: box map in entity form, full Dirichlet mask, continuous
 * right-hand side); the members keep that form (and the gather form
 * without the x-ring), so their single solves give the batch's values. */
int nekbone_batch_get_form(const nekbone_batch *B, int *gather);
/* CG to a relative residual for every problem k: x[k], f[k], r[k] as for
 * nekbone_cg_tol_d (x, r may be NULL); stops when every problem has met
 * tol or after maxiter iterations.  iters[k], final_rtr[k] may be NULL.
 * Returns NEKBONE_NOT_CONVERGED when some problem did not meet tol. */
int nekbone_batch_cg_tol_d(nekbone_batch *B, double *const *x, const double *const *f,
                           double *const *r, double tol, int maxiter, int precond, int *iters,
                           double *final_rtr);
/* Box factorization of np elements, cubic (src/driver.f:196-247). */
int nekbone_cubic(int np, int *mx, int *my, int *mz);

/* Device buffers owned by the problem (n doubles each), for *_d calls. */
int nekbone_alloc_vec(nekbone_problem *P, double **dptr);
int nekbone_free_vec(nekbone_problem *P, double *dptr);
int nekbone_h2d(nekbone_problem *P, double *dst, const double *src, int64_t count);
int nekbone_d2h(nekbone_problem *P, double *dst, const double *src, int64_t count);
/* Synthetic inputs on the device: U[-1,1) counter-based splitmix64
 * (value k = mix(seed + (k+1)*0x9E3779B97F4A7C15)), and the reference's
 * set_f post-processing f = c * dssum(f) (src/driver.f:96-97). */
int nekbone_rand_uniform_d(nekbone_problem *P, double *dst, int64_t count, uint64_t seed);
int nekbone_rhs_from_raw_d(nekbone_problem *P, double *f);
/* The reference's right-hand side, set_f (src/driver.f:88-101).
 * nekbone_set_f_raw: host only, raw[i-1] = sin(1e9 cos(1e9 (i*i))) for
 * i = 1..n with i*i the wrapping 32-bit product of the reference's
 * default INTEGER, sin/cos of the C library.  nekbone_set_f_d: the raw
 * vector of the problem's n points, then dssum and the product with c,
 * into device vector f (reference point order). */
int nekbone_set_f_raw(int64_t n, double *raw);
int nekbone_set_f_d(nekbone_problem *P, double *f);

/* ---------------------------------------------------------------------- */
/* Measurement                                                              */
/* ---------------------------------------------------------------------- */
/* Per-kernel HIP-event timing on the library stream.  Kernel classes:
 * 0 = element operator (ax local), 1 = gather-scatter, 2 = post (shift +
 * mask + pap), 3 = vector update (+ rtr/rtz), 4 = reductions.  mask bit k
 * enables class k (0 disables).  Launches of enabled classes are
 * bracketed by events without host synchronisation; nekbone_timing_get
 * resolves them and returns the launch count and total milliseconds. */
int nekbone_timing_enable(nekbone_problem *P, int mask);
int nekbone_timing_get(nekbone_problem *P, int kclass, int64_t *launches, double *ms);
int nekbone_timing_reset(nekbone_problem *P);
/* Streaming device copy of `bytes` bytes (half read, half written), reps
 * times: the achievable-bandwidth calibration for roofline figures and
 * the reference point for HBM byte counters.  *ms = mean per copy. */
int nekbone_bw_probe(int64_t bytes, int reps, double *ms);
int nekbone_sync(void);

/* ---------------------------------------------------------------------- */
/* Reference-named Fortran-callable entry points                           */
/* ---------------------------------------------------------------------- */
/* Bind the problem the legacy names act on.  A Fortran host calls this
 * once per mesh after the reference's set-up, in place of the COMMON
 * state (/dimn/, /dxyz/, /mymask/, /comm_handles/) the reference reads.
 * gsh is the handle returned by gs_setup_. */
void nekbone_setup_(const int *nelt, const int *nx1, const int *nelx, const int *nely,
                    const int *nelz, const int32_t *lglel, const double *dxm1, const int *gsh);
/* Select the preconditioner solvem_/cg_ apply: 0 = identity (the
 * reference's default build, src/hsmg_dum.f:9-15), nonzero = Jacobi.
 * No counterpart in the reference, which has no Jacobi option.  The
 * choice holds for later meshes; each nekbone_setup_ starts from it, and
 * h1mg_setup_ switches the bound mesh to HSMG. */
void nekbone_use_jacobi_(const int *on);
/* src/cg.f:2-84.  Prints the reference's two residual lines
 * ('cg:',i4,1p4e12.4; src/cg.f:43,76) on stdout.  g and c are uploaded
 * when their address or contents differ from the previous call's. */
void cg_(double *x, const double *f, const double *g, const double *c, double *r, double *w,
         double *p, double *z, const int *n, const int *niter, double *flop_cg);
/* src/semhat.f:2-37: a, c, d (0:n,0:n) column-major, b and z (0:n); w is
 * the reference's work array (unused).  Called by the reference's
 * proxy_setup (src/prox_setup.f:9). */
void semhat_(double *a, double *b, double *c, double *d, double *z, double *w, const int *n);
/* src/cg.f:86-94 (identity, or Jacobi after nekbone_setup_jacobi) */
void solvem_(double *z, double *r, const int *n);
/* src/hsmg.f:3-29; after it solvem_ and cg_ use the HSMG preconditioner,
 * as a build with MGRID does (src/cg.f:86-94) */
void h1mg_setup_(void);
/* src/hsmg.f:960-1034 (rhs is masked in place, as in the reference) */
void h1mg_solve_(double *z, double *rhs, const int *n);
/* src/cg.f:96-125 */
void ax_(double *w, const double *u, const double *gxyz, double *ur, double *us, double *ut,
         double *wk, const int *n);
/* src/cg.f:142-168: one element, g(6,nx1^3) */
void ax_e_(double *w, const double *u, const double *g, double *ur, double *us, double *ut,
           double *wk);
/* src/cg.f:170-187, src/cg.f:189-212 (n = polynomial order) */
void local_grad3_(double *ur, double *us, double *ut, const double *u, const int *n,
                  const double *D, const double *Dt);
void local_grad3_t_(double *u, const double *ur, const double *us, const double *ut, const int *N,
                    const double *D, const double *Dt, double *w);
/* src/prox_dssum.f:2-12 */
void dssum_(double *f);
/* src/cg.f:214-258 (pmask ignored: the mask state lives in the library) */
void maskit_(double *w, double *pmask, const int *nx, const int *ny, const int *nz);
/* src/math.f:603-617 */
double glsc3_(const double *a, const double *b, const double *mult, const int *n);
/* src/jl/gs.c:1160-1170, 1188-1193, 1237-1242 (one rank: comm, np
 * ignored; dom must be 1 = double, op 1 = add) */
void gs_setup_(int *handle, const int64_t *glo_num, const int *n, const int *comm, const int *np);
void gs_op_(const int *handle, double *u, const int *dom, const int *op, const int *transpose);
void gs_free_(const int *handle);

/* ---------------------------------------------------------------------- */
/* Allocation-free entry point.  No counterpart in the reference: it is for */
/* callers that cannot allocate device memory (code under HIP graph         */
/* capture, tight serving loops, harnesses that forbid hipMalloc).          */
/* ---------------------------------------------------------------------- */
/* *result = sum a_i*b_i*mult_i over n points, the value glsc3_ and
 * nekbone_glsc3 return for the same data (bit-identical: same kernels,
 * same block partials).  a, b, mult, result are device pointers; the sum
 * is written to device memory in the library stream's order and the host
 * is not synchronised (nekbone_sync or a stream wait makes it visible).
 * The block partials live in a static device array of 1024 doubles (8 KiB
 * per GPU, held until the process ends), so the call allocates nothing;
 * nekbone_init must have run.  0 <= n <= 2^31 - 1, else NEKBONE_ERR_ARG.
 * Calls on one GPU share its array: calls through the library stream run
 * in order, but two overlapping ones (e.g. replays of a captured call on
 * two streams at once) race. */
int nekbone_glsc3_noalloc_d(const double *a, const double *b, const double *mult, int64_t n,
                            double *result);

#ifdef __cplusplus
}
#endif
#endif
