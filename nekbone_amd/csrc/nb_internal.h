// Internal declarations of the nekbone_amd library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nekbone_amd.h"

namespace nb {

constexpr int kMaxNx = 16;

// Per-point flag byte: Dirichlet list membership, whether the point is
// shared with another element, and its multiplicity (c = 1/multiplicity,
// src/driver.f:274-291) when that fits in 5 bits.
constexpr uint8_t kFlagFull = 1;    // in the full boundary list
constexpr uint8_t kFlagRef = 2;     // in the reference's element-1 list
constexpr uint8_t kFlagShared = 4;  // multiplicity >= 2
constexpr int kMultShift = 3;       // bits 3..7: multiplicity (1..31)

// ---------------------------------------------------------------------
// Errors: the C ABI returns a status and keeps a per-thread message.
// ---------------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define NB_HIP(call)                                                               \
    do {                                                                           \
        hipError_t _e = (call);                                                    \
        if (_e != hipSuccess)                                                      \
            return ::nb::fail(NEKBONE_ERR_HIP, std::string(#call ": ") +           \
                                                   hipGetErrorString(_e));         \
    } while (0)

// The library's stream.  Every device operation of the library, copies and
// fills included, is issued on it.  It is created non-blocking, so the
// legacy null stream (plain hipMemcpy / hipMemset) is not ordered against
// it: a null-stream fill of a new vector may land after a later copy into
// that vector on this stream.
hipStream_t stream();
// Two more streams (k = 0, 1) for work that may run beside the library
// stream's (non-blocking, default priority).  fork_side(k): work queued on
// side stream k from now on follows what stream() has queued so far;
// join_side(k): what stream() queues from now on follows what side stream
// k has queued.  Both are recorded into a graph when stream() is being
// captured.
hipStream_t side_stream(int k);
int fork_side(int k);
int join_side(int k);

// Stream-ordered transfers and fills on stream(); each returns once the
// operation has completed (the host buffer may be reused at once).
int h2d(void *dst, const void *src, size_t bytes);
int d2h(void *dst, const void *src, size_t bytes);
int dzero(void *dst, size_t bytes);

// Device allocation helpers (at least one element is allocated).
template <class T>
int dalloc(T **p, int64_t count)
{
    *p = nullptr;
    if (count <= 0) count = 1;
    NB_HIP(hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * size_t(count)));
    return NEKBONE_OK;
}

// Frees after the library stream has drained, so that no kernel still in
// flight reads or writes the memory.
template <class T>
void dfree(T *&p)
{
    if (p) {
        (void)hipStreamSynchronize(stream());
        (void)hipFree(p);
    }
    p = nullptr;
}

// ---------------------------------------------------------------------
// Host-side set-up (nb_setup.cpp): restatements of the reference's
// set-up algorithms used by the library's own problem constructor.
// ---------------------------------------------------------------------
void gll_points_weights(double *z, double *w, int np);                  // speclib.f ZWGLL
void fd_weights_full(double xx, const double *x, int n, int m, double *c); // semhat.f:39-93
void semhat(double *a, double *b, double *c, double *d, double *z, int n); // semhat.f:2-37
void box_numbering(std::vector<int64_t> &glo, int nx, int nelx, int nely, int nelz,
                   const std::vector<int32_t> &lglel);                   // prox_dssum.f:42-86
// Gather-scatter groups in the jl local order (gs.c:88-113,330-357):
// only ids with >= 2 copies, groups ordered by their smallest index,
// members ascending.  off has ngroups+1 entries.
void gs_groups(const int64_t *glo, int64_t n, std::vector<int32_t> &off,
               std::vector<int32_t> &idx);
// Dirichlet lists (cg.f:214-368): every domain-boundary point, and the
// element-1 positions the reference records on its first maskit call.
void mask_lists(int nx, int nelt, int nelx, int nely, int nelz,
                const std::vector<int32_t> &lglel, std::vector<int32_t> &full,
                std::vector<int32_t> &ref);
// Faces of global element eg (1-based) on the domain boundary, get_face
// (src/cg.f:302-368): x0, x1, y0, y1, z0, z1.
void boundary_faces(int eg, int nelx, int nely, int nelz, bool f[6]);
bool on_face(const bool f[6], int i, int j, int k, int nx);

// Device point order within an element: surface points first, grouped
// so that every shared-point group of a hex mesh touches contiguous runs
// in each element.  Position of reference point (i,j,k), I = nx-2, with
// bc = 0 for an interior coordinate, 1 on the low side, 2 on the high
// side:
//   8 corners            4(bk-1) + 2(bj-1) + (bi-1)
//   12 edges x I         8 + eps*I + (t-1), t the edge's free coordinate;
//                        eps = 2(bj-1)+(bi-1) for the 4 k-edges, then per
//                        k-side 4+4(bk-1) + {bi-1 (j free), 2+bj-1 (i free)}
//   6 faces x I^2        8 + 12I + phi*I^2 + in-face index, phi = x-, x+,
//                        y-, y+, z-, z+; in-face index (k-1)I+(j-1) on x
//                        faces, (k-1)I+(i-1) on y faces, (j-1)I+(i-1) on z
//   interior             nx^3 - I^3 + (k-1)I^2 + (j-1)I + (i-1)
// The element kernel evaluates the same function.
__host__ __device__ inline int surface_order_pos(int nx, int i, int j, int k)
{
    const int I = nx - 2;
    const int bi = i == 0 ? 1 : i == nx - 1 ? 2 : 0;
    const int bj = j == 0 ? 1 : j == nx - 1 ? 2 : 0;
    const int bk = k == 0 ? 1 : k == nx - 1 ? 2 : 0;
    const int nb = (bi != 0) + (bj != 0) + (bk != 0);
    if (nb == 3) return 4 * (bk - 1) + 2 * (bj - 1) + (bi - 1);
    if (nb == 2) {
        int eps, t;
        if (bk == 0) {
            eps = 2 * (bj - 1) + (bi - 1);
            t = k;
        } else if (bj == 0) {
            eps = 4 + 4 * (bk - 1) + (bi - 1);
            t = j;
        } else {
            eps = 4 + 4 * (bk - 1) + 2 + (bj - 1);
            t = i;
        }
        return 8 + eps * I + (t - 1);
    }
    if (nb == 1) {
        int phi, loc;
        if (bi) {
            phi = bi - 1;
            loc = (k - 1) * I + (j - 1);
        } else if (bj) {
            phi = 2 + bj - 1;
            loc = (k - 1) * I + (i - 1);
        } else {
            phi = 4 + bk - 1;
            loc = (j - 1) * I + (i - 1);
        }
        return 8 + 12 * I + phi * I * I + loc;
    }
    return nx * nx * nx - I * I * I + (k - 1) * I * I + (j - 1) * I + (i - 1);
}
// Inverse of surface_order_pos: the reference position i + nx*j + nx^2*k
// of device position d.
__host__ __device__ inline int surface_order_ref(int nx, int d)
{
    const int I = nx - 2, hi = nx - 1;
    int i, j, k;
    if (d < 8) {
        i = (d & 1) ? hi : 0;
        j = (d & 2) ? hi : 0;
        k = (d & 4) ? hi : 0;
    } else if (I > 0 && d < 8 + 12 * I) {
        const int e = (d - 8) / I, t = (d - 8) - e * I + 1;
        if (e < 4) {
            i = (e & 1) ? hi : 0;
            j = (e & 2) ? hi : 0;
            k = t;
        } else {
            const int r = (e - 4) & 3;
            k = ((e - 4) >> 2) ? hi : 0;
            if (r < 2) {
                i = r ? hi : 0;
                j = t;
            } else {
                j = (r - 2) ? hi : 0;
                i = t;
            }
        }
    } else if (I > 0 && d < 8 + 12 * I + 6 * I * I) {
        const int f = d - 8 - 12 * I, phi = f / (I * I), loc = f - phi * (I * I);
        const int u = loc / I + 1, v = loc - (u - 1) * I + 1;
        const int side = (phi & 1) ? hi : 0;
        if (phi < 2) {
            i = side;
            k = u;
            j = v;
        } else if (phi < 4) {
            j = side;
            k = u;
            i = v;
        } else {
            k = side;
            j = u;
            i = v;
        }
    } else {
        const int f = I > 0 ? d - (nx * nx * nx - I * I * I) : 0;
        const int II = I > 0 ? I * I : 1, I1 = I > 0 ? I : 1;
        k = f / II + 1;
        j = (f - (k - 1) * II) / I1 + 1;
        i = f - (k - 1) * II - (j - 1) * I1 + 1;
    }
    return i + nx * (j + nx * k);
}
void boundary_first_perm(int nx, std::vector<int32_t> &perm);
// surface_order_pos is a permutation and surface_order_ref its inverse.
bool surface_order_consistent(int nx);

// ---------------------------------------------------------------------
// Device work of one problem.
// ---------------------------------------------------------------------
// Shared-node groups, split by size so that a wave handles groups of one
// size with a fixed trip count: pairs, quadruples and octuples (faces,
// edges and corners of a hex mesh) as packed member lists, anything else
// as CSR.  Members are device positions in the reference's member order
// (ascending reference local index), which fixes the summation order.
struct GsMap {
    int64_t ngroups = 0, nshared = 0;
    int64_t n2 = 0, n4 = 0, n8 = 0, nr = 0;  // groups per class
    int32_t *d_m2 = nullptr;                  // 2*n2
    int32_t *d_m4 = nullptr;                  // 4*n4
    int32_t *d_m8 = nullptr;                  // 8*n8
    int32_t *d_roff = nullptr;                // nr+1
    int32_t *d_ridx = nullptr;                // members of the other groups
    // Run form of the pair and quad classes: groups that come in runs of
    // r2 (r4) consecutive groups whose every member advances by one device
    // position from group to group (the face and edge interiors of a box
    // mesh in the vertex/edge/face/interior order) are stored as the first
    // group of each run, and group t of a run has members first + t.  The
    // first n2r (n4r) groups of the class are in run form, the rest are in
    // the member lists d_m2 (d_m4).
    int r2 = 0, r4 = 0;
    int64_t n2r = 0, n4r = 0;
    int32_t *d_m2r = nullptr;                 // 2 * n2r / r2
    int32_t *d_m4r = nullptr;                 // 4 * n4r / r4
};

// Per-kernel-class HIP-event timing.  Event pairs are recorded on the
// library stream without synchronising and resolved when read.
// Shape of the element-operator kernel (chosen per order by the
// autotuner): elements per workgroup, geometric factors prefetched at
// kernel entry, x update fused into the element kernel.
struct AxVariant {
    int epb = 1;
    int pf = 0;
    int xu = 0;
};

struct Timing {
    int mask = 0;  // bit k: time kernel class k
    int64_t launches[5] = {0, 0, 0, 0, 0};
    double ms[5] = {0, 0, 0, 0, 0};
    struct Pending {
        hipEvent_t a, b;
        int kclass;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    hipEvent_t cur = nullptr;
};

// Directions kept for the x update of the gather form (x-ring): the
// iteration k stores p_k in ring slot k % kRing and x receives the
// alpha_j p_j of kRing iterations at once (one captured graph chunk).
constexpr int kRing = 10;

// Scalars of a CG solve, resident on the device; the finalize kernels
// update them so that no host round trip is needed per iteration.
struct CgScalars {
    double rtz1, rtz2, beta, pap, alpha, rtr, rtr0, rtz_next, tol2;
    int iter;     // iterations completed
    int done;     // set once rtr <= tol2*rtr0 (tolerance runs)
    int status;   // sticky: 1 = non-finite or non-positive pap / rtz
    int xpending; // x += alpha p of the last completed iteration not yet applied
    double rtr_upd;  // mixed precision: FP64 rtr at the last reliable update
    // x-ring: iterations whose alpha p is in x, and alpha of iteration j
    // in slot j % kRing
    int xdone;
    double aring[kRing];
    // mixed precision, device-driven refreshes (set by k_mx_start): the
    // refresh kernels queued after every chunk of inner iterations run when
    // mx_want holds (inner convergence, residual reduction delta2 since the
    // last refresh, or maxit iterations); fin: the FP64 residual met tol;
    // nupd: refreshes run
    int fin, nupd, maxit;
    double delta2;
};

// Entity table of the gather form of the fused CG (the vector update sums
// the copies of each shared point itself instead of a gather-scatter
// kernel).  Per element kGxStride int32: the element's surface entities in
// device order (8 vertices, 12 edges, 6 faces), each with the device
// position of its first point in every member element of its groups, in
// the groups' member order (the reference's summation order), -1 after
// the last member:
//   vertex v: [8v, 8v+8)   edge k: [64+4k, 68+4k)   face f: [112+2f, 114+2f)
//   [kGxMask]: bit v / 8+k / 20+f set where the entity is on the full
//              Dirichlet list (uniform over the entity's points)
// Point t of an entity has the members slot + t.  Built only when every
// group of the map has this form (box meshes in the device point order).
constexpr int kGxStride = 128;
constexpr int kGxMask = 124;
struct GxView {
    const int32_t *tab;
    int nx, I, n3;
    double inv_n3;   // 1 / n3 (element of a device position)
    float inv_n3f, inv_I, inv_I2;
    int eb;          // elements per workgroup of the per-element update kernels
};
// Elements per workgroup of the gather-form update kernels: about pts
// points per workgroup (measured at 32^3 elements, N=9, on MI355X: 2048
// for FP64, 4096 for FP32), at most kGxMaxEb elements for the table rows
// in LDS.
constexpr int kGxMaxEb = 32;
constexpr int kGxPts64 = 2048, kGxPts32 = 4096;
inline int gx_eb(int nx, int pts)
{
    const int n3 = nx * nx * nx;
    int eb = pts / n3;
    return eb < 1 ? 1 : eb > kGxMaxEb ? kGxMaxEb : eb;
}

// Per-problem arguments of the batched fused CG kernels.
struct CgBatchArgs {
    double *r;
    const double *w, *dinv;
    const uint8_t *flag;
    double *x;
    const double *p;
    int64_t n;
    double *part;       // element + gather-scatter pap parts
    int64_t npap;
    double *part_upd;   // update parts (rtr, then rtz)
    int nupd;
    CgScalars *scal;
    double *hist;
    int hcap;
    // gather form: the problem's entity table view and element count
    GxView G;
    int nelt;
};

// Kernel view of a gather-scatter map: the class member lists and the
// number of workgroups each class takes.
struct GsArgs {
    const int32_t *m2, *m4, *m8, *roff, *ridx;
    int64_t n2, n4, n8, nr;
    int b2, b4, b8, br;  // workgroups per class
    int rev;             // walk the groups downward (see sweep_block)
    const int32_t *m2r, *m4r;  // run form of the pair / quad class (GsMap)
    int r2, r4;
    int64_t n2r, n4r;
    // CG form: partial sums of the element kernel folded into this
    // kernel's (logical workgroup b adds pre[b*kRedThreads + thread]), so
    // the final sum reads this kernel's partials only; null: none
    const double *pre = nullptr;
    int64_t npre = 0;
};
// Device view of a group map (the problem's own: gs_args); the workgroup
// counts b2..br are those of the gather-scatter kernel.
GsArgs gs_args_map(const GsMap &M);
GsArgs gs_args(const nekbone_problem *P);
// Arguments of the fused CG gather-scatter of one problem.
struct GsCgArgs {
    double *w;
    const double *p;
    const uint8_t *flag;
    int mbit;
    GsArgs G;
    double *part;
    const CgScalars *scal;
};


// One level of the HSMG preconditioner (src/hsmg.f).  Level vectors are
// element-blocked nh^3 per element; the finest level is the problem's own
// (device point order, the problem's gather-scatter map), the coarser
// levels use the reference's point order.
struct HsmgLevel {
    int nx = 0, nh = 0, nl = 0;  // order, points per direction, nh+2
    int64_t n = 0;               // nh^3 * nelt
    bool fine = false;
    GsMap gs;                    // coarser levels: own numbering
    int32_t *d_pos = nullptr;    // memory offset of reference point p within an element (nh^3)
    uint8_t *d_shr = nullptr;    // coarser levels: 1 where the point is shared (in a gs group)
    uint8_t *d_bnd = nullptr;    // 1: zeroed by this level's Dirichlet mask
    double *d_c = nullptr;       // inverse multiplicity (restriction weight)
    double *d_sw = nullptr;      // square root of the Schwarz weight (1 away from faces)
    int32_t *d_gsrc = nullptr;   // per element 6*nh^2: memory index of the point the ghost
                                 // receives from the neighbour (-1: none)
    int32_t *d_gin = nullptr;    // per element 6*nh^2: own point one inside that face (memory index)
    int32_t *d_gpart = nullptr;  // per element 6*nh^2: the neighbour's matching ghost slot (-1: none)
    double *d_ghost = nullptr;   // per element 6*nh^2 ghost values of the local solves
    double *d_ops = nullptr;     // distinct 1-D FDM operators: S, S^T (nl*nl each), lambda (nl)
    int nops = 0;
    int16_t *d_eop = nullptr;    // per element: operator index in x, y, z
    double *d_eps = nullptr;     // per element eigenvalue cut-off
    float *d_ops32 = nullptr;    // FP32 copies of d_ops and d_eps (single-precision smoother)
    float *d_eps32 = nullptr;
    double *d_jh = nullptr, *d_jht = nullptr;  // interpolation to the next finer level
    double *d_r = nullptr, *d_e = nullptr, *d_eown = nullptr;  // level work vectors
};

struct Hsmg {
    int lmax = 0;
    HsmgLevel L[4];  // 1..lmax
    double *d_z = nullptr;  // preconditioned residual of the CG (fine level)
    // optional direct coarse solve on the order-1 level (nb_hsmg_coarse.hip)
    int nc = 0;                   // unmasked mesh vertices (coarse unknowns)
    double *d_ainv = nullptr;     // nc x nc inverse of the coarse operator (symmetric)
    int32_t *d_cprim = nullptr;   // nc: a level-1 point holding each unknown
    int32_t *d_cidx = nullptr;    // level-1 points: their unknown, -1 on the mask
    double *d_xc = nullptr;       // nc: coarse solution
    double *d_rc = nullptr;       // nc: coarse right-hand side
};

// Largest coarse problem the direct coarse solve accepts (its dense
// inverse is 512 MB).
constexpr int kMaxCoarse = 8192;

// FP32 copies for the inner iterations of the mixed-precision CG:
// operator data converted from the FP64 originals (rebuilt when the
// geometry or the Jacobi diagonal changes) and the FP32 work vectors,
// all in the device point order.
struct MixedState {
    float *g = nullptr, *Drow = nullptr, *Dcol = nullptr;
    uint16_t *dinv = nullptr;  // Jacobi diagonal as bfloat16 bits
    int mult_in_dinv = 0;      // its last two bits hold log2(multiplicity)
    float *x = nullptr, *r = nullptr, *p = nullptr, *w = nullptr;
    int geom_version = -1, dinv_version = -1;
    AxVariant axv;
    // optional FP16 geometric factors: g = gh * gsc[e][c] (IEEE half bits)
    int g16 = 0;
    uint16_t *gh = nullptr;
    float *gsc = nullptr;
    int g16_version = -1;
    // 1: the FP64 residual refresh flushes x inside the element kernel;
    // 0: in a pass of its own
    int refresh = 1;
};

}  // namespace nb

struct nekbone_problem {
    int nelt = 0, nx1 = 0, nelx = 0, nely = 0, nelz = 0;
    int64_t n = 0;
    unsigned sweep = 0;  // fused CG iterations launched (sets the sweep direction)
    std::vector<int32_t> lglel;
    double dxm1[nb::kMaxNx * nb::kMaxNx] = {};
    double zgll[nb::kMaxNx] = {}, wgll[nb::kMaxNx] = {};
    std::vector<int64_t> glo;
    std::vector<int32_t> mfull, mref;

    // device data
    double *d_D = nullptr;     // dxm1, nx1*nx1 column-major: Dcol[i*nx+m] = D(m,i)
    double *d_Drow = nullptr;  // row-major copy: Drow[i*nx+m] = D(i,m)
    double *d_g = nullptr;     // geometric factors, six planes g[c][e][p] (c*n + e*nxyz + p)
    double *d_c = nullptr;     // inverse multiplicity
    double *d_dinv = nullptr;  // Jacobi
    uint8_t *d_flag = nullptr; // per-point flags (kFlag*), device order
    // Device point order: every per-point vector on the device (x, r, p,
    // w, f, c, dinv, flag) stores reference point p of element e at
    // e*nx^3 + perm[p]; host-facing entry points convert.  g keeps the
    // reference order.
    std::vector<int32_t> perm;
    int32_t *d_perm = nullptr;
    int16_t *d_iref = nullptr;  // inverse: reference position of device position d
    nb::GsMap gs;

    // work vectors (n doubles)
    double *d_x = nullptr, *d_r = nullptr, *d_p = nullptr, *d_w = nullptr, *d_f = nullptr,
           *d_t = nullptr, *d_t2 = nullptr;
    double *d_part = nullptr;  // reduction partials
    int64_t part_cap = 0;
    bool c_is_mult = false;    // c == 1/multiplicity everywhere (fused CG path allowed)
    bool fused = true;         // use the fused CG path when c_is_mult
    int32_t *d_gx = nullptr;   // entity table of the gather form (kGxStride per element)
    int *d_gxbad = nullptr;    // continuity check result (device)
    bool gx_ok = false;        // the map has the entity form (d_gx built)
    int gather = 1;            // fused CG: gather form where available (0: gather-scatter kernel)
    bool gx_cont = true;       // the current solve's right-hand side is continuous (checked)
    double *d_pring = nullptr; // x-ring: kRing direction vectors (n each)
    int xring = 1;             // use the x-ring with the gather form when memory allows
    bool ring_on = false;      // the current solve uses it (set by launch_cg_init)
    nb::AxVariant axv;         // element-kernel shape
    nb::CgScalars *d_scal = nullptr;
    double *d_hist = nullptr;
    int hist_cap = 0;

    int mask_mode = NEKBONE_MASK_REF;
    int mask_built = 0;
    int arith = NEKBONE_ARITH_EXACT;
    bool jacobi = false;
    int geom_version = 0;  // bumped whenever d_g changes
    int dinv_version = 0;  // bumped whenever d_dinv is recomputed
    nb::MixedState *mx = nullptr;
    nb::Hsmg *mg = nullptr;
    int hsmg_fp32 = 0;  // 1: the HSMG local FDM solves run in FP32
    int hsmg_coarse = 0;  // 1: HSMG coarse level solved directly (FP64)
    // captured chunk of CG iterations, replayed while the key matches
    bool use_graph = true;
    hipGraphExec_t graph = nullptr;
    std::vector<uintptr_t> graph_key;
    nb::Timing timing;
};

namespace nb {
inline int check(const nekbone_problem *P)
{
    if (!P) return fail(NEKBONE_ERR_ARG, "null problem");
    return NEKBONE_OK;
}
int ensure_init();
// Mask bit to apply for the next maskit call, advancing the state.
int next_mask_bit(nekbone_problem *P);

// Kernel launchers (nb_kernels.hip / nb_ax.hip).  All on stream().
int ax_blocks(const nekbone_problem *P);
AxVariant ax_default_variant(int nx1);
std::vector<AxVariant> ax_variants(int nx1);
int launch_ax_local(nekbone_problem *P, double *w, const double *u);
// element kernel of the fused CG iteration; z is the vector the CG carries
// (r with the identity, M^-1 r with Jacobi in z form)
// gx: gather form (pap parts over every point, shared ones by the identity
// in nb_ax_kernel.h; the update kernel then sums the shared copies)
// pin/pout (x-ring): the previous direction is read from pin and the new
// one written to pout, and x is not updated; null: P->d_p in place
int launch_ax_cg(nekbone_problem *P, double *x, const double *z, int mbit, int rev = 0,
                 int gx = 0, double *pin = nullptr, double *pout = nullptr);
// x-ring tail before a result is read: x += alpha_j p_j of the iterations
// no flush has folded in yet
int launch_ring_tail(nekbone_problem *P, double *x);
int launch_gs(nekbone_problem *P, double *u, const CgScalars *guard);
// Fused CG variant: finishes shared points (+0.1p, mask) and writes one
// pap part per workgroup (gs_blocks of them) to part.
int launch_gs_cg(nekbone_problem *P, double *w, const double *p, int mbit, double *part,
                 int rev = 0, const double *pre = nullptr, int64_t npre = 0);
// First pass of the mixed solver's FP64 residual refresh (nb_ax.hip):
// x += x32 (+ alpha p32 when pending), x32 = 0, w = A_L x;
// NEKBONE_ERR_UNSUPPORTED when no such kernel is compiled for the shape.
// gate: run only when mx_want holds (device-driven refreshes)
int launch_ax_flush(nekbone_problem *P, float *x32, const float *p32, int gate = 0);
// FP32 gather-scatter (mixed-precision inner solve): plain, and the fused
// CG variant (pap parts in FP64).
int launch_gs32(nekbone_problem *P, float *u);
int launch_gs_cg32(nekbone_problem *P, float *w, const float *p, int mbit, double *part,
                   const double *pre = nullptr, int64_t npre = 0);

int gs_blocks(const nekbone_problem *P);
int launch_post(nekbone_problem *P, double *w, const double *p, int mbit, bool pap);
int launch_mask(nekbone_problem *P, double *w, int mbit);
int launch_glsc3(nekbone_problem *P, const double *a, const double *b, const double *m,
                 double *d_out);
// glsc3 of n points with explicit partials: reduce_blocks(n) <= kMaxRedBlocks
// block partials in part, the sum in d_out (both device memory).
int launch_glsc3_raw(const double *a, const double *b, const double *m, int64_t n,
                     double *part, double *d_out);
// Allocation-free entry points (nb_noalloc.hip): per-device addresses of
// their static device scratch, looked up once per device at nekbone_init
// with that device current.
int noalloc_setup(int device);
int launch_brick_g(nekbone_problem *P);
int launch_deformed_g(nekbone_problem *P, const double *d_vert);
int launch_jacobi_local(nekbone_problem *P, double *dl);
int launch_jacobi_finish(nekbone_problem *P, double *d);
int launch_col2(nekbone_problem *P, double *a, const double *b);
int launch_rand(double *dst, int64_t count, uint64_t seed);
int launch_aos_to_soa(nekbone_problem *P, const double *aos, double *soa, int dir);
// Reference order <-> device order of a per-point vector (dst != src).
int launch_to_dev(nekbone_problem *P, double *dst, const double *src);
int launch_to_ref(nekbone_problem *P, double *dst, const double *src);
// Host vector (reference order) <-> device-order array, through d_t.
int upload_dev(nekbone_problem *P, double *dst, const double *src);
int download_ref(nekbone_problem *P, double *dst, const double *src);
// The scalars of P queued for the host: scal_post copies them into pinned
// slot 0 or 1 after the work queued so far, scal_wait blocks until that
// copy has landed (so the host can check one chunk of iterations while the
// next one runs).
int scal_post(const nekbone_problem *P, int slot);
int scal_wait(int slot, CgScalars &s);
// CG building blocks
// zform (fused Jacobi CG only, see cg_zform): r receives z = M^-1 r;
// fp64_iter: FP64 CG iterations follow (false for the mixed-precision
// solver, whose inner iterations never use the x-ring)
int launch_cg_init(nekbone_problem *P, double *x, double *r, const double *f, int mbit,
                   int precond, double tol2, bool zform = false, bool fp64_iter = true);
// whether the fused Jacobi CG runs in z form (the vector r holds z = M^-1 r
// between launch_cg_init and launch_cg_finish)
bool cg_zform(const nekbone_problem *P, int precond);
// k: the iteration's index within the solve (its x-ring slot)
int launch_cg_iteration(nekbone_problem *P, double *x, double *r, int precond, int mbit, int k);
// x-ring buffers of the problem (FP64) allocated when the device has room
// for them; false: the solve keeps the per-iteration x update
bool ensure_ring(nekbone_problem *P);
// after the last iteration: the pending x update, and r = z / dinv in z form
int launch_cg_finish(nekbone_problem *P, double *x, double *r, bool zform);
int64_t cg_partials_needed(const nekbone_problem *P);
// Single-block scalar kernels of the fused iteration, shared with the
// mixed-precision solver: alpha from npart pap parts; rtr / next rtz from
// 2*nu parts (rtr parts, then rtz parts).
int launch_pap_final(nekbone_problem *P, const double *part, int64_t npart);
int launch_update_final(nekbone_problem *P, const double *part, int nu, int precond,
                        int set_xpending, int xmark = 0);
int cg_update_blocks(const nekbone_problem *P);
// Whether the fused iterations run in gather form: the entity table exists,
// the caller has not switched it off, the full Dirichlet mask applies
// (the element kernel's pap identity needs p zero wherever w is masked)
// and c = 1/multiplicity with the fused path on.
bool cg_gather(const nekbone_problem *P);
// pts: points per workgroup of the update kernel (kGxPts64 / kGxPts32)
GxView gx_view(const nekbone_problem *P, int pts);
// Before a solve (launch_cg_init calls it): whether the masked
// right-hand side f (device order) has equal values at all copies of
// every shared point, which the gather form's pap identity needs (p stays
// continuous then); sets gx_cont.  Synchronises the stream.  A no-op
// returning OK when the gather form does not apply anyway.
int cg_gather_check(nekbone_problem *P, const double *f);
// Right-hand side for the autotuners' timed iterations: uniform random,
// summed over shared copies (continuous, so the gather form applies where
// it would in a solve).
int tune_rhs(nekbone_problem *P, uint64_t seed);
// batched fused CG (nb_batch.hip drives these)
GsCgArgs gs_cg_args(nekbone_problem *P, double *w, const double *p, int mbit, double *part);
int launch_gs_cg_batch(const GsCgArgs *d_args, const int *d_start, int nprob, int nblocks);
// gx: the gather form (no gather-scatter kernel; see cg_gather)
CgBatchArgs cg_batch_args(nekbone_problem *P, double *x, double *r, bool gx = false);
int launch_cg_batch_tail(const CgBatchArgs *d_args, const int *d_ustart, int nprob, int nupd_total,
                         int precond, bool fma, bool xp, int set_xpending, bool gx = false);
int cg_ax_blocks(const nekbone_problem *P);
// whether a fused iteration folds the element kernel's partial sums into
// the gather-scatter kernel's (it has workgroups enough)
bool cg_fold_partials(const nekbone_problem *P);
int cg_gs_blocks(const nekbone_problem *P);
void free_mixed(nekbone_problem *P);
// gather-scatter with an explicit map (HSMG levels)
// guard (may be null): the kernel returns at once when guard->done is set
int launch_gs_map(const GsMap &M, double *u, const CgScalars *guard = nullptr);
int gs_map_upload(GsMap &G, const std::vector<int32_t> &off, const std::vector<int32_t> &idx,
                  const int32_t *pos, int n3);
void gs_map_free(GsMap &G);
// Destroys the problem's captured CG-iteration graph (after the stream has
// drained).  Every routine that frees or replaces a device buffer that the
// iteration's kernels read calls it first: the graph's key holds buffer
// addresses, and a new buffer can come back at a freed one's address.
void drop_problem_graph(nekbone_problem *P);
// HSMG (nb_hsmg_setup.cpp, nb_hsmg.hip)
int hsmg_build(nekbone_problem *P);
void hsmg_free(nekbone_problem *P);
int hsmg_coarse_build(nekbone_problem *P);
void hsmg_coarse_free(Hsmg *M);
int hsmg_coarse_apply(nekbone_problem *P, double *e, const CgScalars *guard = nullptr);
// guard (may be null): every kernel of the apply returns at once when
// guard->done is set (iterations queued past a CG's convergence);
// rtz_part (may be null): the finest interpolation also writes the
// per-element parts of glsc3(r, c, z), nelt of them
int launch_hsmg_apply(nekbone_problem *P, double *z, double *r,
                      const CgScalars *guard = nullptr, double *rtz_part = nullptr);
// One CG iteration with an externally computed z = M^-1 r (r as the
// preconditioner left it): rtz, beta, p, w = A p, pap, alpha, x, r, rtr;
// k: the iteration's index within the solve (its x-ring slot);
// rtz_parts > 0: the rtz parts are already in P->d_part (that many, from
// the preconditioner's last kernel), else they are formed here.
int launch_cg_iteration_z(nekbone_problem *P, double *x, double *r, const double *z, int mbit,
                          int k, int rtz_parts = 0);
int set_c_host(nekbone_problem *P, const double *c);
int timed_begin(nekbone_problem *P, int kclass);
int timed_end(nekbone_problem *P, int kclass);
int timing_resolve(nekbone_problem *P);
int reduce_blocks(int64_t n);
}  // namespace nb
