// Batched CG: many independent problems of different sizes and orders
// solved together.  The reference's driver runs one mesh at a time
// (src/driver.f:28-79); here an iteration of the whole batch is
//   * the element kernel: one persistent launch over a device queue of
//     (problem, element tile) items for the orders 6, 8, 10, 12 (the work-
//     queue kernel, k_ax_queue), or one launch per order group; other
//     orders always launch per group;
//   * one gather-scatter launch for all problems, or none in the gather
//     form (cg_gather), which the batch takes when every member can;
//   * one launch each for all problems of pap/alpha, the vector update with
//     rtr and rtz (in the gather form also the sums of the shared copies),
//     and rtr/beta.
// The autotuner picks the element-kernel path, the tile shapes per order
// and the x-update placement.  Each problem keeps its own workgroups,
// partial sums and device scalars, so every value is the one its single-
// problem solve produces with the same shape.  Chunks of iterations are
// captured once into a HIP graph and replayed; a problem that has
// converged keeps its state (its kernels return at once) until the whole
// batch is done.
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
        int nblk_ax = 0, threads = 0;
        nb::axk::AxArgs *d_ax = nullptr;
        int *d_axstart = nullptr;
    };
    std::vector<nekbone_problem *> P;
    std::vector<Group> groups;
    int xu = 1;
    bool fma = false;
    nb::CgBatchArgs *d_args = nullptr;
    int *d_ustart = nullptr;
    int nupd_total = 0;
    // work-queue element kernel: index of its compiled tile set, -1 for
    // per-group launches
    int queue = -1;
    nb::axk::AxArgs *d_qargs = nullptr;
    int *d_qnx = nullptr;
    nb::axk::QItem *d_qitems = nullptr;
    int nqitems = 0, qgrid = 0, qprobs = 0, qitems_cap = 0;
    size_t qlds = 0;
    // gather-scatter of every problem in one launch
    nb::GsCgArgs *d_gsall = nullptr;
    int *d_gsstart = nullptr;
    int nblk_gs = 0;
    // captured chunk of iterations; shapes counts changes of the shapes
    hipGraphExec_t exec = nullptr;
    int exec_precond = -1, exec_chunk = 0, exec_shapes = -1;
    int shapes = 0;
    int use_graph = 1;  // replay chunks of iterations from a captured graph
    // gather form for every member (the iteration has no gather-scatter
    // kernel); decided per solve, see decide_form
    bool gx = false;
    int exec_gx = -1;
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

const std::vector<axk::QShape> &tableq()
{
    static std::vector<axk::QShape> t;
    static bool done = false;
    if (!done) {
        done = true;
        axk::regq_all(t);
    }
    return t;
}

// slot of an order in the work-queue kernel's tile set, -1 if it has none
int qslot(int nx1)
{
    for (int i = 0; i < 4; ++i)
        if (axk::kQOrders[i] == nx1) return i;
    return -1;
}

bool in_queue(const nekbone_batch *B, int nx1) { return B->queue >= 0 && qslot(nx1) >= 0; }

// Elements per workgroup of every member as the current path runs it; the
// problems carry the same shape, so a single solve gives the same values.
void apply_shapes(nekbone_batch *B)
{
    for (auto &G : B->groups) {
        const int epb = in_queue(B, G.nx1) ? tableq()[size_t(B->queue)].e[qslot(G.nx1)] : G.v.epb;
        for (int k : G.members) B->P[size_t(k)]->axv = AxVariant{epb, in_queue(B, G.nx1) ? 0 : G.v.pf, B->xu};
    }
    ++B->shapes;
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

// Before the members' CG init: every member may take the gather form
// (without the x-ring, so that a member's single solve runs the kernels
// the batch runs).
void offer_gather(nekbone_batch *B)
{
    for (nekbone_problem *P : B->P) {
        P->gather = 1;
        if (P->xring) {
            P->xring = 0;
            drop_problem_graph(P);
        }
    }
}

// After the members' CG init (which checks each right-hand side): the
// batch runs the gather form when every member can; otherwise the
// gather-scatter form, and the members keep that setting, so that their
// single solves give the batch's values.
void decide_form(nekbone_batch *B)
{
    bool all = true;
    for (nekbone_problem *P : B->P) all = all && cg_gather(P);
    B->gx = all;
    if (!all)
        for (nekbone_problem *P : B->P) P->gather = 0;
}

void drop_graph(nekbone_batch *B)
{
    if (B->exec) {
        // no replay of it may still be running
        (void)hipStreamSynchronize(stream());
        (void)hipGraphExecDestroy(B->exec);
    }
    B->exec = nullptr;
    B->exec_precond = -1;
}

// Element-kernel arguments of problem k for the CG mode.
axk::AxArgs ax_args(nekbone_batch *B, int k, int mbit)
{
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
    A.mbit = mbit;
    A.scal = P->d_scal;
    A.part = P->d_part;
    A.nelt = P->nelt;
    A.gx = B->gx ? 1 : 0;
    return A;
}

int device_cus()
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
            return 0;
        cus = prop.multiProcessorCount;
    }
    return cus;
}

// Argument tables of one solve: the element-kernel arguments of the work
// queue and of the per-group launches, the gather-scatter arguments of
// every problem, and per problem the scalar and update arguments.  The
// members' shapes (apply_shapes) must be current.
int prepare(nekbone_batch *B, int precond, const std::vector<int> &mbit)
{
    (void)precond;
    // Jacobi runs in z form (see cg_zform): the element kernel's identity
    // mode reads z
    const int mode = 1;
    for (auto &G : B->groups) {
        G.fn = nullptr;
        if (in_queue(B, G.nx1)) continue;
        auto it = tableb().find(axk::Key(G.nx1, G.v.epb, B->fma ? 1 : 0, mode, G.v.pf, B->xu));
        if (it == tableb().end())
            return fail(NEKBONE_ERR_UNSUPPORTED, "batch: no batched kernel for this order");
        G.fn = it->second;
        std::vector<axk::AxArgs> ax;
        std::vector<int> axs{0};
        for (int k : G.members) {
            ax.push_back(ax_args(B, k, mbit[size_t(k)]));
            axs.push_back(axs.back() + cg_ax_blocks(B->P[size_t(k)]));
        }
        G.nblk_ax = axs.back();
        G.threads = axk::round64(G.nx1 * G.nx1 * G.v.epb);
        if (int rc = upload(&G.d_ax, ax)) return rc;
        if (int rc = upload(&G.d_axstart, axs)) return rc;
    }
    B->nqitems = 0;
    if (B->queue >= 0) {
        const axk::QShape &Q = tableq()[size_t(B->queue)];
        std::vector<axk::AxArgs> qa;
        std::vector<int> qnx;
        std::vector<axk::QItem> items;
        size_t lds = 0;
        for (int slot = 3; slot >= 0; --slot)  // highest order first
            for (auto &G : B->groups) {
                if (G.nx1 != axk::kQOrders[slot]) continue;
                const size_t n3 = size_t(G.nx1) * G.nx1 * G.nx1;
                lds = std::max(lds, 2 * size_t(Q.e[slot]) * n3 * sizeof(double));
                for (int k : G.members) {
                    const int q = int(qa.size());
                    qa.push_back(ax_args(B, k, mbit[size_t(k)]));
                    qnx.push_back(G.nx1);
                    for (int b = 0; b < cg_ax_blocks(B->P[size_t(k)]); ++b)
                        items.push_back(axk::QItem{q, b});
                }
            }
        // the argument tables keep their size for the life of the batch
        // (the queue's membership is fixed); the item list is reallocated
        // when a tile set changes the number of workgroups (the captured
        // graph is re-captured then, see apply_shapes)
        if (B->d_qitems && int(items.size()) != B->qitems_cap) {
            dfree(B->d_qitems);
            B->qitems_cap = 0;
        }
        if (int rc = upload(&B->d_qargs, qa)) return rc;
        if (int rc = upload(&B->d_qnx, qnx)) return rc;
        if (int rc = upload(&B->d_qitems, items)) return rc;
        B->qitems_cap = int(items.size());
        B->qprobs = int(qa.size());
        B->nqitems = int(items.size());
        B->qlds = lds;
        int per_cu = 0;
        NB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(Q.fn[B->fma ? 1 : 0][B->xu]), Q.ntb, lds));
        const int cus = device_cus();
        if (per_cu < 1 || cus < 1)
            return fail(NEKBONE_ERR_UNSUPPORTED, "batch: work-queue kernel does not fit a CU");
        B->qgrid = std::max(1, std::min(B->nqitems, per_cu * cus));
    }
    std::vector<GsCgArgs> gs;
    std::vector<int> gss{0};
    for (size_t k = 0; k < B->P.size() && !B->gx; ++k) {
        nekbone_problem *P = B->P[k];
        gs.push_back(gs_cg_args(P, P->d_w, P->d_p, mbit[k], P->d_part + cg_ax_blocks(P)));
        if (cg_fold_partials(P)) {  // as in the single-problem iteration
            gs.back().G.pre = P->d_part;
            gs.back().G.npre = cg_ax_blocks(P);
        }
        gss.push_back(gss.back() + cg_gs_blocks(P));
    }
    B->nblk_gs = gss.back();
    if (!B->gx) {
        if (int rc = upload(&B->d_gsall, gs)) return rc;
        if (int rc = upload(&B->d_gsstart, gss)) return rc;
    }
    std::vector<CgBatchArgs> args;
    std::vector<int> us{0};
    for (nekbone_problem *P : B->P) {
        args.push_back(cg_batch_args(P, P->d_x, P->d_r, B->gx));
        us.push_back(us.back() + args.back().nupd);
    }

    B->nupd_total = us.back();
    if (int rc = upload(&B->d_args, args)) return rc;
    return upload(&B->d_ustart, us);
}

int launch_iteration(nekbone_batch *B, int precond)
{
    const int np = int(B->P.size());
    if (B->queue >= 0 && B->nqitems > 0) {
        const axk::QShape &Q = tableq()[size_t(B->queue)];
        hipLaunchKernelGGL(Q.fn[B->fma ? 1 : 0][B->xu], dim3(B->qgrid), dim3(Q.ntb), B->qlds,
                           stream(), static_cast<const axk::AxArgs *>(B->d_qargs),
                           static_cast<const int *>(B->d_qnx),
                           static_cast<const axk::QItem *>(B->d_qitems), B->nqitems);
        NB_HIP(hipGetLastError());
    }
    // the order groups' element kernels write disjoint problems: they run
    // on the library stream and the two side streams at once, so one
    // group's tail overlaps another's launch; each group goes, largest
    // first, to the stream with the fewest points queued so far
    std::vector<std::pair<int64_t, size_t>> order;
    for (size_t g = 0; g < B->groups.size(); ++g) {
        const auto &G = B->groups[g];
        if (!G.fn) continue;
        int64_t pts = 0;
        for (int k : G.members) pts += B->P[size_t(k)]->n;
        order.emplace_back(-pts, g);
    }
    std::sort(order.begin(), order.end());
    const int nside = std::min(int(order.size()) - 1, 2);
    for (int k = 0; k < nside; ++k)
        if (int rc = fork_side(k)) return rc;
    int64_t load[3] = {0, 0, 0};
    for (const auto &og : order) {
        int lane = 0;
        for (int l = 1; l <= nside; ++l)
            if (load[l] < load[lane]) lane = l;
        load[lane] -= og.first;
        const auto &G = B->groups[og.second];
        hipStream_t st = lane == 0 ? stream() : side_stream(lane - 1);
        hipLaunchKernelGGL(G.fn, dim3(G.nblk_ax), dim3(G.threads), 0, st,
                           static_cast<const axk::AxArgs *>(G.d_ax),
                           static_cast<const int *>(G.d_axstart), int(G.members.size()));
        NB_HIP(hipGetLastError());
    }
    for (int k = 0; k < nside; ++k)
        if (int rc = join_side(k)) return rc;
    if (!B->gx)
        if (int rc = launch_gs_cg_batch(B->d_gsall, B->d_gsstart, np, B->nblk_gs)) return rc;
    return launch_cg_batch_tail(B->d_args, B->d_ustart, np, B->nupd_total, precond, B->fma,
                                !B->xu, B->xu, B->gx);
}

int ensure_graph(nekbone_batch *B, int precond, int chunk)
{
    if (B->exec && B->exec_precond == precond && B->exec_chunk == chunk &&
        B->exec_shapes == B->shapes && B->exec_gx == int(B->gx))
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
    B->exec_shapes = B->shapes;
    B->exec_gx = int(B->gx);
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
    // per-group launches at the orders' default shapes (measured faster
    // than the work queue on the batched workload; the autotuner compares
    // both); every member runs the path's shape, in the batch and alone
    apply_shapes(B);
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
    }
    dfree(B->d_qargs);
    dfree(B->d_qnx);
    dfree(B->d_qitems);
    dfree(B->d_gsall);
    dfree(B->d_gsstart);
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
    apply_shapes(B);
    return NEKBONE_OK;
}

extern "C" int nekbone_batch_set_queue(nekbone_batch *B, int tileset)
{
    if (!B) return fail(NEKBONE_ERR_ARG, "batch: null");
    if (tileset < -1 || tileset >= int(tableq().size()))
        return fail(NEKBONE_ERR_ARG, "batch: tile set out of range");
    B->queue = tileset;
    apply_shapes(B);
    return NEKBONE_OK;
}

extern "C" int nekbone_batch_get_path(const nekbone_batch *B, int *tileset, int *xu, int *ntilesets)
{
    if (!B) return fail(NEKBONE_ERR_ARG, "batch: null");
    if (tileset) *tileset = B->queue;
    if (xu) *xu = B->xu;
    if (ntilesets) *ntilesets = int(tableq().size());
    return NEKBONE_OK;
}

extern "C" int nekbone_batch_get_form(const nekbone_batch *B, int *gather)
{
    if (!B) return fail(NEKBONE_ERR_ARG, "batch: null");
    if (gather) *gather = B->gx ? 1 : 0;
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
    offer_gather(B);
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
    decide_form(B);
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

// Batch autotuner: times batched iterations of every element-kernel path
// (per-group launches at the orders' default shapes and at doubled ones,
// the work-queue kernel with each compiled tile set) with the x update
// carried by the element kernel or by the vector update; keeps the fastest
// and gives the members its shapes.
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
    struct Cand {
        int queue, dbl, xu, pf;
    };
    std::vector<Cand> cands;
    bool any_q = false;
    for (auto &G : B->groups) any_q = any_q || qslot(G.nx1) >= 0;
    for (int xu = 1; xu >= 0; --xu) {
        cands.push_back(Cand{-1, 0, xu, 0});
        cands.push_back(Cand{-1, 0, xu, 3});
        cands.push_back(Cand{-1, 1, xu, 0});
        if (any_q)
            for (int q = 0; q < int(tableq().size()); ++q) cands.push_back(Cand{q, 0, xu, 0});
    }
    // per-group shape of a candidate: the order's default, the default with
    // the factors loaded one t-step ahead, or twice as many elements per
    // workgroup, where that batched kernel is compiled
    auto set_cand = [&](const Cand &c) {
        B->queue = c.queue;
        B->xu = c.xu;
        const int fm = B->fma ? 1 : 0;
        for (auto &G : B->groups) {
            const int e = axk::epb_for(G.nx1);
            const bool has2 = tableb().count(axk::Key(G.nx1, 2 * e, fm, 1, 0, c.xu)) != 0;
            const bool has3 = tableb().count(axk::Key(G.nx1, e, fm, 1, 3, c.xu)) != 0;
            if (c.dbl && has2)
                G.v = AxVariant{2 * e, 0, 1};
            else
                G.v = AxVariant{e, c.pf == 3 && has3 ? 3 : 0, 1};
        }
        apply_shapes(B);
    };
    hipEvent_t e0, e1;
    NB_HIP(hipEventCreate(&e0));
    NB_HIP(hipEventCreate(&e1));
    double best = 1e300;
    Cand win{B->queue, 0, B->xu, 0};
    bool have_win = false;
    int rc = NEKBONE_OK;
    for (const Cand &c : cands) {
        set_cand(c);
        offer_gather(B);
        for (size_t k = 0; k < np && !rc; ++k) {
            nekbone_problem *P = B->P[k];
            // continuous right-hand sides, as a solve's (the form is timed
            // that the solves will run)
            rc = tune_rhs(P, 12345 + k);
            if (!rc)
                rc = launch_cg_init(P, P->d_x, P->d_r, P->d_f, 1, precond, 0.0,
                                    precond == NEKBONE_PRECOND_JACOBI);
        }
        decide_form(B);
        if (!rc) rc = prepare(B, precond, mbit);
        for (int i = 0; i < 2 && !rc; ++i) rc = launch_iteration(B, precond);
        if (!rc && hipEventRecord(e0, stream()) != hipSuccess) rc = NEKBONE_ERR_HIP;
        for (int i = 0; i < 6 && !rc; ++i) rc = launch_iteration(B, precond);
        if (!rc && hipEventRecord(e1, stream()) != hipSuccess) rc = NEKBONE_ERR_HIP;
        if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = NEKBONE_ERR_HIP;
        float ms = 0.f;
        if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = NEKBONE_ERR_HIP;
        if (rc) break;
        if (ms / 6 < best) {
            best = ms / 6;
            win = c;
            have_win = true;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (have_win) set_cand(win);
    drop_graph(B);
    if (rc) return rc == NEKBONE_ERR_HIP ? fail(rc, "batch autotune: HIP error") : rc;
    if (best_ms) *best_ms = best;
    return NEKBONE_OK;
}
