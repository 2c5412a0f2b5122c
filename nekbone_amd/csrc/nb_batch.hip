// Batched CG: many independent problems of different sizes and orders
// solved together.  The reference's driver runs one mesh at a time
// (src/driver.f:28-79); here every iteration of every problem in the batch
// is one launch per kernel per order group (element kernel, gather-scatter)
// plus one launch for all problems (pap/alpha, r update with rtr and rtz,
// rtr/beta).  Each problem keeps its own workgroups, partial sums and
// device scalars, so every value is the one its single-problem solve
// produces.  Chunks of iterations are captured once into a HIP graph and
// replayed; a problem that has converged keeps its state (its kernels
// return at once) until the whole batch is done.
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "nb_ax_kernel.h"

struct nekbone_batch {
    struct Group {
        int nx1 = 0;
        nb::AxVariant v;
        std::vector<int> members;
        nb::axk::AxFnB fn = nullptr;
        int nblk_ax = 0, nblk_gs = 0, threads = 0;
        nb::axk::AxArgs *d_ax = nullptr;
        int *d_axstart = nullptr;
        nb::GsCgArgs *d_gs = nullptr;
        int *d_gsstart = nullptr;
    };
    std::vector<nekbone_problem *> P;
    std::vector<Group> groups;
    int xu = 1;
    bool fma = false;
    nb::CgBatchArgs *d_args = nullptr;
    int *d_ustart = nullptr;
    int nupd_total = 0;
    // captured chunk of iterations
    hipGraphExec_t exec = nullptr;
    int exec_precond = -1, exec_chunk = 0, exec_xu = -1;
    int use_graph = 1;  // replay chunks of iterations from a captured graph
};

namespace nb {
namespace {

axk::TableB &tableb()
{
    static axk::TableB t;
    static bool done = false;
    if (!done) {
        done = true;
        axk::regb_all(t);
    }
    return t;
}

// Argument tables keep their device address for the life of the batch (a
// captured graph refers to it); their sizes depend only on the membership,
// which is fixed at creation, so later solves rewrite them in place.
template <class T>
int upload(T **d, const std::vector<T> &h)
{
    if (!*d)
        if (int rc = dalloc(d, int64_t(h.size()))) return rc;
    // stream-ordered: kernels of an earlier replay that read the table
    // finish before it is rewritten
    return h2d(*d, h.data(), sizeof(T) * h.size());
}

void drop_graph(nekbone_batch *B)
{
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    B->exec = nullptr;
    B->exec_precond = -1;
}

// Argument tables of one solve: per group the element-kernel and
// gather-scatter arguments of its members, per problem the scalar and
// update arguments.
int prepare(nekbone_batch *B, int precond, const std::vector<int> &mbit)
{
    // Jacobi runs in z form (see cg_zform): the element kernel's identity
    // mode reads z
    const int mode = 1;
    for (auto &G : B->groups) {
        G.fn = nullptr;
        auto it = tableb().find(axk::Key(G.nx1, G.v.epb, B->fma ? 1 : 0, mode, G.v.pf, B->xu));
        if (it == tableb().end())
            return fail(NEKBONE_ERR_UNSUPPORTED, "batch: no batched kernel for this order");
        G.fn = it->second;
        std::vector<axk::AxArgs> ax;
        std::vector<GsCgArgs> gs;
        std::vector<int> axs{0}, gss{0};
        for (int k : G.members) {
            nekbone_problem *P = B->P[size_t(k)];
            axk::AxArgs A{};
            A.w = P->d_w;
            A.p = P->d_p;
            A.r = P->d_r;
            A.x = P->d_x;
            A.g = P->d_g;
            A.gcs = P->n;
            A.Drow = P->d_Drow;
            A.Dcol = P->d_D;
            A.flag = P->d_flag;
            A.iref = P->d_iref;
            A.mbit = mbit[size_t(k)];
            A.scal = P->d_scal;
            A.part = P->d_part;
            A.nelt = P->nelt;
            ax.push_back(A);
            axs.push_back(axs.back() + cg_ax_blocks(P));
            gs.push_back(gs_cg_args(P, P->d_w, P->d_p, mbit[size_t(k)],
                                    P->d_part + cg_ax_blocks(P)));
            if (cg_fold_partials(P)) {  // as in the single-problem iteration
                gs.back().G.pre = P->d_part;
                gs.back().G.npre = cg_ax_blocks(P);
            }
            gss.push_back(gss.back() + cg_gs_blocks(P));
        }
        G.nblk_ax = axs.back();
        G.nblk_gs = gss.back();
        G.threads = axk::round64(G.nx1 * G.nx1 * G.v.epb);
        if (int rc = upload(&G.d_ax, ax)) return rc;
        if (int rc = upload(&G.d_axstart, axs)) return rc;
        if (int rc = upload(&G.d_gs, gs)) return rc;
        if (int rc = upload(&G.d_gsstart, gss)) return rc;
    }
    std::vector<CgBatchArgs> args;
    std::vector<int> us{0};
    for (nekbone_problem *P : B->P) {
        args.push_back(cg_batch_args(P, P->d_x, P->d_r));
        us.push_back(us.back() + args.back().nupd);
    }
    B->nupd_total = us.back();
    if (int rc = upload(&B->d_args, args)) return rc;
    return upload(&B->d_ustart, us);
}

int launch_iteration(nekbone_batch *B, int precond)
{
    const int np = int(B->P.size());
    for (auto &G : B->groups) {
        hipLaunchKernelGGL(G.fn, dim3(G.nblk_ax), dim3(G.threads), 0, stream(),
                           static_cast<const axk::AxArgs *>(G.d_ax),
                           static_cast<const int *>(G.d_axstart), int(G.members.size()));
        NB_HIP(hipGetLastError());
        if (int rc = launch_gs_cg_batch(G.d_gs, G.d_gsstart, int(G.members.size()), G.nblk_gs))
            return rc;
    }
    return launch_cg_batch_tail(B->d_args, B->d_ustart, np, B->nupd_total, precond, B->fma,
                                !B->xu, B->xu);
}

int ensure_graph(nekbone_batch *B, int precond, int chunk)
{
    if (B->exec && B->exec_precond == precond && B->exec_chunk == chunk && B->exec_xu == B->xu)
        return NEKBONE_OK;
    drop_graph(B);
    hipGraph_t g = nullptr;
    NB_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
    int rc = NEKBONE_OK;
    for (int i = 0; i < chunk && !rc; ++i) rc = launch_iteration(B, precond);
    hipError_t e = hipStreamEndCapture(stream(), &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    NB_HIP(e);
    e = hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    NB_HIP(e);
    B->exec_precond = precond;
    B->exec_chunk = chunk;
    B->exec_xu = B->xu;
    return NEKBONE_OK;
}

int read_done(nekbone_batch *B, std::vector<CgScalars> &s)
{
    s.resize(B->P.size());
    for (size_t k = 0; k < B->P.size(); ++k)
        NB_HIP(hipMemcpyAsync(&s[k], B->P[k]->d_scal, sizeof(CgScalars), hipMemcpyDeviceToHost,
                              stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

}  // namespace
}  // namespace nb

using namespace nb;

// cubic, src/driver.f:196-247: the box mx x my x mz = np elements with the
// smallest ratio of its largest to its smallest side among the
// factorizations the reference searches.
extern "C" int nekbone_cubic(int np, int *mx, int *my, int *mz)
{
    if (np < 1 || !mx || !my || !mz) return fail(NEKBONE_ERR_ARG, "cubic: np >= 1");
    *mx = np;
    *my = 1;
    *mz = 1;
    double ratio = double(np);
    const int iroot3 = int(std::pow(double(np), 1. / 3.) + 0.000001);
    int ix = 0, iy = 0;
    for (int i = iroot3; i >= 1; --i) {
        int iz = i;
        const int myx = np / iz;
        if (np - myx * iz != 0) continue;
        const int iroot2 = int(std::pow(double(myx), 1. / 2.) + 0.000001);
        for (int j = iroot2; j >= 1; --j) {
            iy = j;
            ix = myx / iy;
            if (myx - ix * iy == 0) break;
        }
        if (ix < iy) std::swap(ix, iy);
        if (ix < iz) std::swap(ix, iz);
        if (iy < iz) std::swap(iy, iz);
        if (double(ix) / iz < ratio) {
            ratio = double(ix) / iz;
            *mx = ix;
            *my = iy;
            *mz = iz;
        }
    }
    return NEKBONE_OK;
}

extern "C" int nekbone_batch_create(nekbone_batch **out, nekbone_problem *const *probs, int nprob)
{
    if (!out || !probs || nprob < 1) return fail(NEKBONE_ERR_ARG, "batch: need problems");
    *out = nullptr;
    auto *B = new nekbone_batch();
    std::map<int, int> gid;
    for (int k = 0; k < nprob; ++k) {
        nekbone_problem *P = probs[k];
        if (!P) {
            delete B;
            return fail(NEKBONE_ERR_ARG, "batch: null problem");
        }
        if (!P->c_is_mult) {
            delete B;
            return fail(NEKBONE_ERR_UNSUPPORTED, "batch: every problem needs c = 1/multiplicity");
        }
        if (k == 0) B->fma = P->arith == NEKBONE_ARITH_FMA;
        if ((P->arith == NEKBONE_ARITH_FMA) != B->fma) {
            delete B;
            return fail(NEKBONE_ERR_UNSUPPORTED, "batch: problems differ in arithmetic mode");
        }
        B->P.push_back(P);
        auto it = gid.find(P->nx1);
        if (it == gid.end()) {
            gid[P->nx1] = int(B->groups.size());
            nekbone_batch::Group G;
            G.nx1 = P->nx1;
            G.v = AxVariant{axk::epb_for(P->nx1), 0, 1};
            B->groups.push_back(G);
            it = gid.find(P->nx1);
        }
        B->groups[size_t(it->second)].members.push_back(k);
    }
    // every member runs the group's shape, in the batch and alone
    for (auto &G : B->groups)
        for (int k : G.members) B->P[size_t(k)]->axv = AxVariant{G.v.epb, G.v.pf, B->xu};
    *out = B;
    return NEKBONE_OK;
}

extern "C" int nekbone_batch_destroy(nekbone_batch *B)
{
    if (!B) return NEKBONE_OK;
    (void)hipStreamSynchronize(stream());
    drop_graph(B);
    for (auto &G : B->groups) {
        dfree(G.d_ax);
        dfree(G.d_axstart);
        dfree(G.d_gs);
        dfree(G.d_gsstart);
    }
    dfree(B->d_args);
    dfree(B->d_ustart);
    delete B;
    return NEKBONE_OK;
}

extern "C" int nekbone_batch_size(const nekbone_batch *B) { return B ? int(B->P.size()) : -1; }

extern "C" int nekbone_batch_set_graph(nekbone_batch *B, int on)
{
    if (!B) return fail(NEKBONE_ERR_ARG, "batch: null");
    B->use_graph = on ? 1 : 0;
    if (!B->use_graph) drop_graph(B);
    return NEKBONE_OK;
}

extern "C" int nekbone_batch_set_xu(nekbone_batch *B, int xu)
{
    if (!B) return fail(NEKBONE_ERR_ARG, "batch: null");
    B->xu = xu ? 1 : 0;
    for (auto &G : B->groups)
        for (int k : G.members) B->P[size_t(k)]->axv.xu = B->xu;
    return NEKBONE_OK;
}

extern "C" int nekbone_batch_cg_tol_d(nekbone_batch *B, double *const *x, const double *const *f,
                                      double *const *r, double tol, int maxiter, int precond,
                                      int *iters, double *final_rtr)
{
    if (!B) return fail(NEKBONE_ERR_ARG, "batch: null");
    if (!(tol > 0.0)) return fail(NEKBONE_ERR_ARG, "batch: tol must be > 0");
    if (maxiter < 0) return fail(NEKBONE_ERR_ARG, "batch: maxiter < 0");
    if (precond != NEKBONE_PRECOND_JACOBI && precond != NEKBONE_PRECOND_IDENTITY)
        return fail(NEKBONE_ERR_UNSUPPORTED, "batch: identity or Jacobi preconditioner");
    const size_t np = B->P.size();
    std::vector<int> mbit(np);
    for (size_t k = 0; k < np; ++k) {
        nekbone_problem *P = B->P[k];
        if (precond == NEKBONE_PRECOND_JACOBI && !P->jacobi)
            return fail(NEKBONE_ERR_STATE, "batch: Jacobi not set up on every problem");
        if (!f || !f[k]) return fail(NEKBONE_ERR_ARG, "batch: null right-hand side");
        if (int rc = launch_to_dev(P, P->d_f, f[k])) return rc;
        const int mbit0 = next_mask_bit(P);
        if (int rc = launch_cg_init(P, P->d_x, P->d_r, P->d_f, mbit0, precond, tol * tol,
                                    precond == NEKBONE_PRECOND_JACOBI))
            return rc;
        mbit[k] = next_mask_bit(P);
    }
    if (int rc = prepare(B, precond, mbit)) return rc;
    const int chunk = 10;
    if (B->use_graph)
        if (int rc = ensure_graph(B, precond, chunk)) return rc;
    std::vector<CgScalars> s;
    int done_iters = 0;
    while (done_iters < maxiter) {
        if (B->use_graph && maxiter - done_iters >= chunk) {
            NB_HIP(hipGraphLaunch(B->exec, stream()));
            done_iters += chunk;
        } else if (maxiter - done_iters >= chunk) {
            for (int i = 0; i < chunk; ++i)
                if (int rc = launch_iteration(B, precond)) return rc;
            done_iters += chunk;
        } else {
            for (; done_iters < maxiter; ++done_iters)
                if (int rc = launch_iteration(B, precond)) return rc;
        }
        if (int rc = read_done(B, s)) return rc;
        bool all = true;
        for (const CgScalars &q : s) all = all && q.done;
        if (all) break;
    }
    for (size_t k = 0; k < np; ++k) {
        nekbone_problem *P = B->P[k];
        if (int rc = launch_cg_finish(P, P->d_x, P->d_r, precond == NEKBONE_PRECOND_JACOBI))
            return rc;
        if (x && x[k])
            if (int rc = launch_to_ref(P, x[k], P->d_x)) return rc;
        if (r && r[k])
            if (int rc = launch_to_ref(P, r[k], P->d_r)) return rc;
    }
    if (int rc = read_done(B, s)) return rc;
    int status = 0;
    bool all = true;
    for (size_t k = 0; k < np; ++k) {
        if (iters) iters[k] = s[k].iter;
        if (final_rtr) final_rtr[k] = s[k].rtr;
        status |= s[k].status;
        all = all && s[k].done;
    }
    if (status) return fail(NEKBONE_ERR_NUMERIC, "batch: non-positive or non-finite pap");
    if (!all) return fail(NEKBONE_NOT_CONVERGED, "batch: a problem stopped at maxiter unconverged");
    return NEKBONE_OK;
}

// Batch autotuner: the element kernels of every order group with the x
// update carried by the element kernel or by the vector update, timed over
// batched iterations; keeps the faster.
extern "C" int nekbone_batch_autotune(nekbone_batch *B, int precond, double *best_ms)
{
    if (!B) return fail(NEKBONE_ERR_ARG, "batch: null");
    if (precond != NEKBONE_PRECOND_JACOBI && precond != NEKBONE_PRECOND_IDENTITY)
        return fail(NEKBONE_ERR_UNSUPPORTED, "batch: identity or Jacobi preconditioner");
    const size_t np = B->P.size();
    std::vector<int> mbit(np, 1);
    for (size_t k = 0; k < np; ++k) {
        nekbone_problem *P = B->P[k];
        if (precond == NEKBONE_PRECOND_JACOBI && !P->jacobi)
            return fail(NEKBONE_ERR_STATE, "batch: Jacobi not set up on every problem");
        mbit[k] = P->mask_mode == NEKBONE_MASK_DIRICHLET ? 1 : 2;
    }
    hipEvent_t e0, e1;
    NB_HIP(hipEventCreate(&e0));
    NB_HIP(hipEventCreate(&e1));
    double best = 1e300;
    int win = B->xu, rc = NEKBONE_OK;
    for (int xu = 0; xu <= 1 && !rc; ++xu) {
        nekbone_batch_set_xu(B, xu);
        for (size_t k = 0; k < np && !rc; ++k) {
            nekbone_problem *P = B->P[k];
            rc = launch_rand(P->d_f, P->n, 12345 + k);
            if (!rc)
                rc = launch_cg_init(P, P->d_x, P->d_r, P->d_f, 1, precond, 0.0,
                                    precond == NEKBONE_PRECOND_JACOBI);
        }
        if (!rc) rc = prepare(B, precond, mbit);
        for (int i = 0; i < 2 && !rc; ++i) rc = launch_iteration(B, precond);
        if (!rc && hipEventRecord(e0, stream()) != hipSuccess) rc = NEKBONE_ERR_HIP;
        for (int i = 0; i < 6 && !rc; ++i) rc = launch_iteration(B, precond);
        if (!rc && hipEventRecord(e1, stream()) != hipSuccess) rc = NEKBONE_ERR_HIP;
        if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = NEKBONE_ERR_HIP;
        float ms = 0.f;
        if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = NEKBONE_ERR_HIP;
        if (!rc && ms / 6 < best) {
            best = ms / 6;
            win = xu;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    nekbone_batch_set_xu(B, win);
    drop_graph(B);
    if (rc) return rc == NEKBONE_ERR_HIP ? fail(rc, "batch autotune: HIP error") : rc;
    if (best_ms) *best_ms = best;
    return NEKBONE_OK;
}
