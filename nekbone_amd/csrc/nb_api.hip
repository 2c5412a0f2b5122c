// C ABI of nekbone_amd: problem lifecycle, operators, CG drivers, timing,
// and the reference-named Fortran-callable entry points.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <thread>
#include <tuple>
#include <vector>

#include "nb_internal.h"

namespace nb {

static thread_local std::string t_err;
void set_error(const std::string &msg) { t_err = msg; }
int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

namespace {
int g_device = -1;
hipStream_t g_stream = nullptr;
// two more streams for work that runs beside the library stream's
// (created like it: non-blocking, default priority), and the events that
// fork work to them and join it back
constexpr int kSides = 2;
hipStream_t g_side[kSides] = {nullptr, nullptr};
hipEvent_t g_fork[kSides] = {nullptr, nullptr}, g_join[kSides] = {nullptr, nullptr};

__global__ void k_fill(double *a, double v, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_recip(double *a, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 1. / a[i];
}
int grid(int64_t n) { return int((n + 255) / 256); }

int fill(double *a, double v, int64_t n)
{
    if (n <= 0) return NEKBONE_OK;
    hipLaunchKernelGGL(k_fill, dim3(grid(n)), dim3(256), 0, stream(), a, v, n);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

std::map<std::tuple<int, int, int>, AxVariant> &tuned_cache()
{
    static std::map<std::tuple<int, int, int>, AxVariant> m;
    return m;
}

// A shape found by nekbone_autotune for this order (any preconditioner),
// else the built-in default.
AxVariant tuned_or_default(int nx1)
{
    auto it = tuned_cache().find(std::make_tuple(nx1, -1, -1));
    return it != tuned_cache().end() ? it->second : ax_default_variant(nx1);
}

}  // namespace

hipStream_t stream() { return g_stream; }
hipStream_t side_stream(int k) { return g_side[k]; }

int fork_side(int k)
{
    NB_HIP(hipEventRecord(g_fork[k], g_stream));
    NB_HIP(hipStreamWaitEvent(g_side[k], g_fork[k], 0));
    return NEKBONE_OK;
}

int join_side(int k)
{
    NB_HIP(hipEventRecord(g_join[k], g_side[k]));
    NB_HIP(hipStreamWaitEvent(g_stream, g_join[k], 0));
    return NEKBONE_OK;
}

namespace {
CgScalars *g_hscal = nullptr;  // pinned: two slots of posted scalars
hipEvent_t g_sev[2] = {nullptr, nullptr};
int g_sev_dev = -1;
}  // namespace

int scal_post(const nekbone_problem *P, int slot)
{
    if (g_sev_dev != g_device) {
        if (!g_hscal)
            NB_HIP(hipHostMalloc(reinterpret_cast<void **>(&g_hscal), 2 * sizeof(CgScalars),
                                 hipHostMallocDefault));
        for (hipEvent_t &e : g_sev) {
            if (e) (void)hipEventDestroy(e);
            e = nullptr;
            NB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        g_sev_dev = g_device;
    }
    NB_HIP(hipMemcpyAsync(&g_hscal[slot], P->d_scal, sizeof(CgScalars), hipMemcpyDeviceToHost,
                          stream()));
    NB_HIP(hipEventRecord(g_sev[slot], stream()));
    return NEKBONE_OK;
}

int scal_wait(int slot, CgScalars &s)
{
    NB_HIP(hipEventSynchronize(g_sev[slot]));
    s = g_hscal[slot];
    return NEKBONE_OK;
}

int ensure_init()
{
    if (g_stream) return NEKBONE_OK;
    return nekbone_init(g_device < 0 ? 0 : g_device);
}

int h2d(void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return NEKBONE_OK;
    if (int rc = ensure_init()) return rc;
    NB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

int d2h(void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return NEKBONE_OK;
    if (int rc = ensure_init()) return rc;
    NB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

int dzero(void *dst, size_t bytes)
{
    if (bytes == 0) return NEKBONE_OK;
    if (int rc = ensure_init()) return rc;
    NB_HIP(hipMemsetAsync(dst, 0, bytes, stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

int next_mask_bit(nekbone_problem *P)
{
    if (P->mask_mode == NEKBONE_MASK_DIRICHLET) return 1;
    if (!P->mask_built) {
        P->mask_built = 1;
        return 1;
    }
    return 2;
}

static int take_event(Timing &T, hipEvent_t *ev)
{
    if (!T.pool.empty()) {
        *ev = T.pool.back();
        T.pool.pop_back();
        return NEKBONE_OK;
    }
    NB_HIP(hipEventCreate(ev));
    return NEKBONE_OK;
}

int timed_begin(nekbone_problem *P, int kclass)
{
    Timing &T = P->timing;
    if (!(T.mask >> kclass & 1)) return NEKBONE_OK;
    if (int rc = take_event(T, &T.cur)) return rc;
    NB_HIP(hipEventRecord(T.cur, stream()));
    return NEKBONE_OK;
}

int timed_end(nekbone_problem *P, int kclass)
{
    Timing &T = P->timing;
    if (!(T.mask >> kclass & 1) || !T.cur) return NEKBONE_OK;
    hipEvent_t b;
    if (int rc = take_event(T, &b)) return rc;
    NB_HIP(hipEventRecord(b, stream()));
    T.pending.push_back({T.cur, b, kclass});
    T.cur = nullptr;
    if (T.pending.size() >= 4096) return timing_resolve(P);
    return NEKBONE_OK;
}

int timing_resolve(nekbone_problem *P)
{
    Timing &T = P->timing;
    if (T.pending.empty()) return NEKBONE_OK;
    NB_HIP(hipEventSynchronize(T.pending.back().b));
    for (const Timing::Pending &q : T.pending) {
        float ms = 0.f;
        NB_HIP(hipEventElapsedTime(&ms, q.a, q.b));
        T.launches[q.kclass] += 1;
        T.ms[q.kclass] += double(ms);
        T.pool.push_back(q.a);
        T.pool.push_back(q.b);
    }
    T.pending.clear();
    return NEKBONE_OK;
}

// Point flags: list membership, sharing, multiplicity (see kFlag*).
static void build_flags(nekbone_problem *P, const std::vector<int32_t> &off,
                        const std::vector<int32_t> &idx, std::vector<uint8_t> &flag,
                        std::vector<int32_t> &mult)
{
    const int64_t n = P->n;
    flag.assign(size_t(n), 0);
    mult.assign(size_t(n), 1);
    for (size_t g = 0; g + 1 < off.size(); ++g)
        for (int32_t q = off[g]; q < off[g + 1]; ++q) mult[size_t(idx[size_t(q)])] = off[g + 1] - off[g];
    for (int32_t l : P->mfull) flag[size_t(l)] |= kFlagFull;
    for (int32_t l : P->mref) flag[size_t(l)] |= kFlagRef;
    for (int64_t l = 0; l < n; ++l) {
        const int m = mult[size_t(l)];
        if (m >= 2) flag[size_t(l)] |= kFlagShared;
        if (m < 32) flag[size_t(l)] |= uint8_t(m << kMultShift);
    }
}

// Whether c equals 1/multiplicity at every point (what the reference's
// set_multiplicity produces); then the fused CG path may derive c from the
// flag byte instead of streaming the c array.
static bool c_matches_mult(const double *c, const std::vector<int32_t> &mult)
{
    for (size_t l = 0; l < mult.size(); ++l)
        if (mult[l] >= 32 || c[l] != 1. / double(mult[l])) return false;
    return true;
}

// Device position of reference local index l.
static int64_t dev_index(const nekbone_problem *P, int64_t l)
{
    const int64_t n3 = int64_t(P->nx1) * P->nx1 * P->nx1;
    const int64_t e = l / n3;
    return e * n3 + P->perm[size_t(l - e * n3)];
}

// Upload a host vector in reference order into a device-order array.
int upload_dev(nekbone_problem *P, double *dst, const double *src)
{
    NB_HIP(hipMemcpyAsync(P->d_t, src, sizeof(double) * size_t(P->n), hipMemcpyHostToDevice,
                          stream()));
    return launch_to_dev(P, dst, P->d_t);
}

// Download a device-order array into a host vector in reference order.
int download_ref(nekbone_problem *P, double *dst, const double *src)
{
    if (int rc = launch_to_ref(P, P->d_t, src)) return rc;
    NB_HIP(hipMemcpyAsync(dst, P->d_t, sizeof(double) * size_t(P->n), hipMemcpyDeviceToHost,
                          stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

// Entity table of the gather form of the fused CG (GxView): from the
// groups (members as reference local indices, in the reference's member
// order) and the device-order flags.  False when some group does not have
// the entity form (all members at the same offset within an entity of the
// same kind, one slot list per entity) or some entity is only partly on
// the Dirichlet list; the fused CG then keeps the gather-scatter kernel.
static bool build_gx_table(const nekbone_problem *P, const std::vector<int32_t> &off,
                           const std::vector<int32_t> &idx,
                           const std::vector<uint8_t> &flag_dev, std::vector<int32_t> &tab)
{
    const int nx = P->nx1, I = nx - 2, n3 = nx * nx * nx;
    const int64_t nelt = P->nelt;
    tab.assign(size_t(nelt) * kGxStride, -1);
    for (int64_t e = 0; e < nelt; ++e)
        for (int k = kGxMask; k < kGxStride; ++k) tab[size_t(e * kGxStride + k)] = 0;
    struct Ent {
        int sb, K, o, bit;
    };
    // entity of device position d within an element (false: interior)
    auto ent = [&](int d, Ent &x) -> bool {
        if (d < 8) {
            x = Ent{8 * d, 8, 0, d};
            return true;
        }
        if (I > 0 && d < 8 + 12 * I) {
            const int k = (d - 8) / I;
            x = Ent{64 + 4 * k, 4, (d - 8) - k * I, 8 + k};
            return true;
        }
        if (I > 0 && d < 8 + 12 * I + 6 * I * I) {
            const int f = (d - 8 - 12 * I) / (I * I);
            x = Ent{112 + 2 * f, 2, (d - 8 - 12 * I) - f * I * I, 20 + f};
            return true;
        }
        return false;
    };
    std::vector<int64_t> pos;
    for (size_t g = 0; g + 1 < off.size(); ++g) {
        const int S = off[g + 1] - off[g];
        pos.clear();
        int o0 = -1;
        for (int32_t q = off[g]; q < off[g + 1]; ++q) {
            const int64_t D = dev_index(P, idx[size_t(q)]);
            const int64_t e = D / n3;
            Ent x;
            if (!ent(int(D - e * n3), x) || S > x.K) return false;
            if (q == off[g])
                o0 = x.o;
            else if (x.o != o0)
                return false;
            pos.push_back(D);
        }
        for (int k = 0; k < S; ++k) {
            const int64_t e = pos[size_t(k)] / n3;
            Ent x;
            ent(int(pos[size_t(k)] - e * n3), x);
            int32_t *sl = &tab[size_t(e * kGxStride + x.sb)];
            if (sl[0] == -1) {
                for (int m = 0; m < S; ++m) sl[m] = int32_t(pos[size_t(m)] - o0);
            } else {
                for (int m = 0; m < x.K; ++m)
                    if (sl[m] != (m < S ? int32_t(pos[size_t(m)] - o0) : -1)) return false;
            }
        }
    }
    int8_t seen[26];
    for (int64_t e = 0; e < nelt; ++e) {
        std::fill(seen, seen + 26, int8_t(-1));
        for (int d = 0; d < n3; ++d) {
            Ent x;
            if (!ent(d, x)) continue;
            const int8_t m = (flag_dev[size_t(e * n3 + d)] & kFlagFull) ? 1 : 0;
            if (seen[x.bit] < 0) {
                seen[x.bit] = m;
                if (m) tab[size_t(e * kGxStride + kGxMask)] |= int32_t(1u << x.bit);
            } else if (seen[x.bit] != m) {
                return false;
            }
        }
    }
    return true;
}

int set_c_host(nekbone_problem *P, const double *c)
{
    std::vector<int32_t> off, idx;
    gs_groups(P->glo.data(), P->n, off, idx);
    std::vector<uint8_t> flag;
    std::vector<int32_t> mult;
    build_flags(P, off, idx, flag, mult);
    if (int rc = upload_dev(P, P->d_c, c)) return rc;
    NB_HIP(hipStreamSynchronize(stream()));
    P->c_is_mult = c_matches_mult(c, mult);
    return NEKBONE_OK;
}

// Upload the set-up data common to both constructors and allocate the
// work space.  Expects n, nelt, nx1, dxm1, lglel, glo, masks filled in.
static int finish_problem(nekbone_problem *P, const double *c_host)
{
    const int64_t n = P->n;
    const int nx = P->nx1;
    if (!surface_order_consistent(nx))
        return fail(NEKBONE_ERR_STATE, "internal: device point order is not a permutation");
    boundary_first_perm(nx, P->perm);
    if (int rc = dalloc(&P->d_perm, int64_t(P->perm.size()))) return rc;
    if (int rc = h2d(P->d_perm, P->perm.data(), sizeof(int32_t) * P->perm.size())) return rc;
    {
        std::vector<int16_t> iref(P->perm.size());
        for (size_t p = 0; p < P->perm.size(); ++p) iref[size_t(P->perm[p])] = int16_t(p);
        if (int rc = dalloc(&P->d_iref, int64_t(iref.size()))) return rc;
        if (int rc = h2d(P->d_iref, iref.data(), sizeof(int16_t) * iref.size())) return rc;
    }

    // Groups keep the reference's member order (ascending reference local
    // index, which fixes the summation order); members are stored as
    // device positions.
    std::vector<int32_t> off, idx;
    gs_groups(P->glo.data(), n, off, idx);
    P->gs.ngroups = int64_t(off.size()) - 1;
    P->gs.nshared = int64_t(idx.size());
    std::vector<int32_t> m2, m4, m8, roff{0}, ridx;
    for (size_t g = 0; g + 1 < off.size(); ++g) {
        const int sz = off[g + 1] - off[g];
        std::vector<int32_t> &dst = sz == 2 ? m2 : sz == 4 ? m4 : sz == 8 ? m8 : ridx;
        for (int32_t q = off[g]; q < off[g + 1]; ++q)
            dst.push_back(int32_t(dev_index(P, idx[size_t(q)])));
        if (sz != 2 && sz != 4 && sz != 8) roff.push_back(int32_t(ridx.size()));
    }
    // Within each fixed-size class the groups are ordered by the device
    // position of their first member (the member order inside a group,
    // which fixes the sum, is untouched).  On a box mesh a wave of pair
    // groups then covers one face of an element, whose points sit
    // contiguously in device order, against the matching contiguous face
    // of the neighbour; edge and corner groups likewise.
    auto order_class = [](std::vector<int32_t> &m, int S) {
        const size_t ng = m.size() / size_t(S);
        std::vector<size_t> ord(ng);
        for (size_t g = 0; g < ng; ++g) ord[g] = g;
        std::stable_sort(ord.begin(), ord.end(),
                         [&](size_t a, size_t b) { return m[a * S] < m[b * S]; });
        std::vector<int32_t> out(m.size());
        for (size_t g = 0; g < ng; ++g)
            std::copy(m.begin() + ptrdiff_t(ord[g] * S), m.begin() + ptrdiff_t(ord[g] * S + S),
                      out.begin() + ptrdiff_t(g * S));
        m.swap(out);
    };
    order_class(m2, 2);
    order_class(m4, 4);
    order_class(m8, 8);
    GsMap &G = P->gs;
    G.n2 = int64_t(m2.size()) / 2;
    G.n4 = int64_t(m4.size()) / 4;
    G.n8 = int64_t(m8.size()) / 8;
    G.nr = int64_t(roff.size()) - 1;
    auto up = [](int32_t **d, const std::vector<int32_t> &h) -> int {
        if (int rc = dalloc(d, int64_t(h.size()))) return rc;
        return h2d(*d, h.data(), sizeof(int32_t) * h.size());
    };
    for (auto pr : {std::make_pair(&G.d_m2, &m2), std::make_pair(&G.d_m4, &m4),
                    std::make_pair(&G.d_m8, &m8), std::make_pair(&G.d_roff, &roff),
                    std::make_pair(&G.d_ridx, &ridx)})
        if (int rc = up(pr.first, *pr.second)) return rc;
    // Run form of the pair and quad classes (face and edge interiors of a
    // box: runs of (nx-2)^2 and nx-2 groups).  The class's groups are split
    // into maximal runs (every member one position past the previous
    // group's); whole multiples of R go to the run form, placed first,
    // the remainders stay in the member list.
    auto split = [](std::vector<int32_t> &m, int S, int R, std::vector<int32_t> &runs) {
        runs.clear();
        std::vector<int32_t> rest;
        const size_t ng = m.size() / size_t(S);
        int64_t nrun = 0;
        for (size_t g = 0; g < ng;) {
            size_t len = 1;
            while (g + len < ng) {
                bool next = true;
                for (int k = 0; k < S; ++k)
                    next = next && m[(g + len) * S + k] == m[(g + len - 1) * S + k] + 1;
                if (!next) break;
                ++len;
            }
            const size_t full = R >= 2 ? len / size_t(R) * size_t(R) : 0;
            for (size_t c = 0; c < full; c += size_t(R))
                for (int k = 0; k < S; ++k) runs.push_back(m[(g + c) * S + k]);
            for (size_t t = full; t < len; ++t)
                for (int k = 0; k < S; ++k) rest.push_back(m[(g + t) * S + k]);
            nrun += int64_t(full);
            g += len;
        }
        m.swap(rest);
        return nrun;
    };
    {
        std::vector<int32_t> r2, r4;
        const int R2 = (nx - 2) * (nx - 2), R4 = nx - 2;
        const int64_t n2r = split(m2, 2, R2, r2), n4r = split(m4, 4, R4, r4);
        dfree(G.d_m2);
        dfree(G.d_m4);
        if (int rc = up(&G.d_m2, m2)) return rc;
        if (int rc = up(&G.d_m4, m4)) return rc;
        if (n2r > 0) {
            if (int rc = up(&G.d_m2r, r2)) return rc;
            G.r2 = R2;
            G.n2r = n2r;
        }
        if (n4r > 0) {
            if (int rc = up(&G.d_m4r, r4)) return rc;
            G.r4 = R4;
            G.n4r = n4r;
        }
    }

    std::vector<double> drow(size_t(nx) * nx);
    for (int i = 0; i < nx; ++i)
        for (int m = 0; m < nx; ++m) drow[size_t(i * nx + m)] = P->dxm1[i + nx * m];
    if (int rc = dalloc(&P->d_D, int64_t(nx) * nx)) return rc;
    if (int rc = dalloc(&P->d_Drow, int64_t(nx) * nx)) return rc;
    if (int rc = h2d(P->d_D, P->dxm1, sizeof(double) * nx * nx)) return rc;
    if (int rc = h2d(P->d_Drow, drow.data(), sizeof(double) * nx * nx)) return rc;
    if (int rc = dalloc(&P->d_g, 6 * n)) return rc;
    if (int rc = dalloc(&P->d_c, n)) return rc;
    for (double **v : {&P->d_x, &P->d_r, &P->d_p, &P->d_w, &P->d_f, &P->d_t, &P->d_t2})
        if (int rc = dalloc(v, n)) return rc;
    P->part_cap = cg_partials_needed(P);
    if (int rc = dalloc(&P->d_part, P->part_cap)) return rc;
    if (int rc = dalloc(&P->d_scal, 1)) return rc;
    if (int rc = dzero(P->d_scal, sizeof(CgScalars))) return rc;
    if (int rc = dzero(P->d_part, sizeof(double) * size_t(P->part_cap))) return rc;
    for (double *v : {P->d_x, P->d_r, P->d_p, P->d_w, P->d_f, P->d_t, P->d_t2})
        if (int rc = dzero(v, sizeof(double) * size_t(n))) return rc;

    std::vector<uint8_t> flag, flag_dev(static_cast<size_t>(n));
    std::vector<int32_t> mult;
    build_flags(P, off, idx, flag, mult);
    for (int64_t l = 0; l < n; ++l) flag_dev[size_t(dev_index(P, l))] = flag[size_t(l)];
    if (int rc = dalloc(&P->d_flag, n)) return rc;
    if (int rc = h2d(P->d_flag, flag_dev.data(), size_t(n))) return rc;
    {
        std::vector<int32_t> tab;
        if (build_gx_table(P, off, idx, flag_dev, tab)) {
            if (int rc = dalloc(&P->d_gx, int64_t(tab.size()))) return rc;
            if (int rc = h2d(P->d_gx, tab.data(), sizeof(int32_t) * tab.size())) return rc;
            if (int rc = dalloc(&P->d_gxbad, 1)) return rc;
            P->gx_ok = true;
        }
    }

    if (c_host) {
        if (int rc = upload_dev(P, P->d_c, c_host)) return rc;
        P->c_is_mult = c_matches_mult(c_host, mult);
    } else {
        // set_multiplicity, src/driver.f:274-291: c = 1 / QQ^T(1)
        if (int rc = fill(P->d_c, 1.0, n)) return rc;
        if (int rc = launch_gs(P, P->d_c, nullptr)) return rc;
        hipLaunchKernelGGL(k_recip, dim3(grid(n)), dim3(256), 0, stream(), P->d_c, n);
        NB_HIP(hipGetLastError());
        bool ok = true;
        for (int32_t m : mult) ok = ok && m < 32;
        P->c_is_mult = ok;
    }
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

static int ensure_hist(nekbone_problem *P, int niter)
{
    const int need = 1 + 3 * std::max(niter, 0) + 3;
    if (P->hist_cap >= need) return NEKBONE_OK;
    drop_problem_graph(P);
    dfree(P->d_hist);
    if (int rc = dalloc(&P->d_hist, need)) return rc;
    P->hist_cap = need;
    return NEKBONE_OK;
}

static int check_precond(nekbone_problem *P, int precond)
{
    if (precond == NEKBONE_PRECOND_JACOBI && !P->jacobi)
        return fail(NEKBONE_ERR_STATE, "cg: Jacobi preconditioner not set up");
    if (precond == NEKBONE_PRECOND_HSMG && !P->mg)
        return fail(NEKBONE_ERR_STATE, "cg: HSMG preconditioner not set up");
    if (precond != NEKBONE_PRECOND_JACOBI && precond != NEKBONE_PRECOND_IDENTITY &&
        precond != NEKBONE_PRECOND_HSMG)
        return fail(NEKBONE_ERR_UNSUPPORTED, "cg: unknown preconditioner");
    return NEKBONE_OK;
}

// One iteration: solveM (src/cg.f:86-94) is fused into the iteration for
// the identity and Jacobi; HSMG runs h1mg_solve first (which masks r, as
// the reference's does) and the iteration takes its z.
static int cg_step(nekbone_problem *P, double *x, double *r, int precond, int mbit, int k)
{
    if (precond == NEKBONE_PRECOND_HSMG) {
        // applies queued past convergence return at once; the finest
        // interpolation leaves the rtz parts, one per element
        if (int rc = launch_hsmg_apply(P, P->mg->d_z, r, P->d_scal, P->d_part)) return rc;
        return launch_cg_iteration_z(P, x, r, P->mg->d_z, mbit, k, P->nelt);
    }
    return launch_cg_iteration(P, x, r, precond, mbit, k);
}

void drop_problem_graph(nekbone_problem *P)
{
    if (P->graph) {
        // no replay of it may still be running
        (void)hipStreamSynchronize(stream());
        (void)hipGraphExecDestroy(P->graph);
    }
    P->graph = nullptr;
    P->graph_key.clear();
}

// Iterations it0..it0+count-1, from a captured graph of kGraphChunk
// iterations when possible.  Launch arguments of an iteration depend only
// on the problem's buffers and settings (all scalars live on the device)
// and, with the x-ring, on the iteration's index modulo kRing, so a chunk
// captured once replays any later chunk that starts at a multiple of
// kGraphChunk, of any solve with the same key; event timing is not
// captured (it needs host bookkeeping).
constexpr int kGraphChunk = 10;
static_assert(kGraphChunk % kRing == 0, "a captured chunk covers whole x-ring turns");
static int run_iterations(nekbone_problem *P, double *x, double *r, int precond, int mbit,
                          int count, int it0 = 0)
{
    const bool graphable = P->use_graph && P->timing.mask == 0 && it0 % kGraphChunk == 0;
    int done = 0;
    if (graphable && count >= kGraphChunk) {
        const std::vector<uintptr_t> key = {
            uintptr_t(x), uintptr_t(r), uintptr_t(precond), uintptr_t(mbit),
            uintptr_t(P->d_hist), uintptr_t(P->hist_cap), uintptr_t(P->axv.epb),
            uintptr_t(P->axv.pf), uintptr_t(P->axv.xu), uintptr_t(P->arith),
            uintptr_t(P->fused), uintptr_t(cg_gather(P)), uintptr_t(P->mg), uintptr_t(P->d_dinv),
            uintptr_t(P->hsmg_fp32), uintptr_t(P->hsmg_coarse), uintptr_t(P->ring_on),
            uintptr_t(P->d_pring)};
        if (!P->graph || P->graph_key != key) {
            drop_problem_graph(P);
            hipGraph_t g = nullptr;
            NB_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
            int rc = NEKBONE_OK;
            for (int i = 0; i < kGraphChunk && !rc; ++i) rc = cg_step(P, x, r, precond, mbit, i);
            hipError_t e = hipStreamEndCapture(stream(), &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            NB_HIP(e);
            e = hipGraphInstantiate(&P->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            NB_HIP(e);
            P->graph_key = key;
        }
        for (; done + kGraphChunk <= count; done += kGraphChunk)
            NB_HIP(hipGraphLaunch(P->graph, stream()));
    }
    for (; done < count; ++done)
        if (int rc = cg_step(P, x, r, precond, mbit, it0 + done)) return rc;
    return NEKBONE_OK;
}

// CG on device vectors (src/cg.f:2-84).
static int run_cg(nekbone_problem *P, double *x, const double *f, double *r, int niter,
                  int precond, double tol2)
{
    if (int rc = check_precond(P, precond)) return rc;
    if (niter < 0) return fail(NEKBONE_ERR_ARG, "cg: niter < 0");
    if (int rc = ensure_hist(P, niter)) return rc;
    const int mbit0 = next_mask_bit(P);  // maskit(r), src/cg.f:40
    if (int rc = launch_cg_init(P, x, r, f, mbit0, precond, tol2, cg_zform(P, precond)))
        return rc;
    const int mbit = next_mask_bit(P);  // maskit in ax, src/cg.f:119 (state is now fixed)
    if (int rc = run_iterations(P, x, r, precond, mbit, niter)) return rc;
    return launch_cg_finish(P, x, r, cg_zform(P, precond));
}

}  // namespace nb

using namespace nb;

// ======================================================================
// Lifecycle
// ======================================================================
extern "C" const char *nekbone_last_error(void) { return t_err.c_str(); }

extern "C" int nekbone_init(int device)
{
    if (g_stream && device == g_device) return NEKBONE_OK;
    int count = 0;
    NB_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(NEKBONE_ERR_ARG, "nekbone_init: no such device");
    NB_HIP(hipSetDevice(device));
    if (int rc = noalloc_setup(device)) return rc;
    if (g_stream) NB_HIP(hipStreamDestroy(g_stream));
    for (int k = 0; k < kSides; ++k) {
        if (g_side[k]) NB_HIP(hipStreamDestroy(g_side[k]));
        if (g_fork[k]) NB_HIP(hipEventDestroy(g_fork[k]));
        if (g_join[k]) NB_HIP(hipEventDestroy(g_join[k]));
    }
    NB_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    for (int k = 0; k < kSides; ++k) {
        NB_HIP(hipStreamCreateWithFlags(&g_side[k], hipStreamNonBlocking));
        NB_HIP(hipEventCreateWithFlags(&g_fork[k], hipEventDisableTiming));
        NB_HIP(hipEventCreateWithFlags(&g_join[k], hipEventDisableTiming));
    }
    g_device = device;
    return NEKBONE_OK;
}

extern "C" int nekbone_finalize(void)
{
    if (g_stream) {
        NB_HIP(hipStreamSynchronize(g_stream));
        NB_HIP(hipStreamDestroy(g_stream));
    }
    for (int k = 0; k < kSides; ++k) {
        if (g_side[k]) {
            NB_HIP(hipStreamSynchronize(g_side[k]));
            NB_HIP(hipStreamDestroy(g_side[k]));
        }
        if (g_fork[k]) NB_HIP(hipEventDestroy(g_fork[k]));
        if (g_join[k]) NB_HIP(hipEventDestroy(g_join[k]));
        g_side[k] = nullptr;
        g_fork[k] = g_join[k] = nullptr;
    }
    g_stream = nullptr;
    return NEKBONE_OK;
}

extern "C" int nekbone_sync(void)
{
    if (int rc = ensure_init()) return rc;
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

// ======================================================================
// Problem set-up
// ======================================================================
extern "C" int nekbone_problem_create_box(nekbone_problem **out, int nelx, int nely, int nelz,
                                          int nx1)
{
    if (!out) return fail(NEKBONE_ERR_ARG, "null output");
    *out = nullptr;
    if (nx1 < 2 || nx1 > kMaxNx || nelx < 1 || nely < 1 || nelz < 1)
        return fail(NEKBONE_ERR_ARG, "create_box: need 2 <= nx1 <= 16 and nel* >= 1");
    const int64_t nelt64 = int64_t(nelx) * nely * nelz;
    const int64_t n64 = nelt64 * nx1 * nx1 * nx1;
    if (n64 >= (int64_t(1) << 31)) return fail(NEKBONE_ERR_ARG, "create_box: n >= 2^31");
    if (int rc = ensure_init()) return rc;
    auto P = std::make_unique<nekbone_problem>();
    P->nelx = nelx;
    P->nely = nely;
    P->nelz = nelz;
    P->nelt = int(nelt64);
    P->nx1 = nx1;
    P->axv = tuned_or_default(nx1);
    P->n = n64;
    P->lglel.resize(size_t(P->nelt));
    for (int e = 0; e < P->nelt; ++e) P->lglel[size_t(e)] = e + 1;
    double a[kMaxNx * kMaxNx], cc[kMaxNx * kMaxNx];
    semhat(a, P->wgll, cc, P->dxm1, P->zgll, nx1 - 1);
    box_numbering(P->glo, nx1, nelx, nely, nelz, P->lglel);
    mask_lists(nx1, P->nelt, nelx, nely, nelz, P->lglel, P->mfull, P->mref);
    if (int rc = finish_problem(P.get(), nullptr)) {
        nekbone_problem_destroy(P.release());
        return rc;
    }
    if (int rc = launch_brick_g(P.get())) {
        nekbone_problem_destroy(P.release());
        return rc;
    }
    NB_HIP(hipStreamSynchronize(stream()));
    *out = P.release();
    return NEKBONE_OK;
}

extern "C" int nekbone_problem_create(nekbone_problem **out, int nelt, int nx1, int nelx,
                                      int nely, int nelz, const int32_t *lglel,
                                      const double *dxm1, const int64_t *glo_num,
                                      const double *gxyz, const double *c)
{
    if (!out) return fail(NEKBONE_ERR_ARG, "null output");
    *out = nullptr;
    if (nx1 < 2 || nx1 > kMaxNx || nelt < 1 || !lglel || !dxm1 || !glo_num)
        return fail(NEKBONE_ERR_ARG, "create: bad arguments");
    if (int64_t(nelx) * nely * nelz < 1) return fail(NEKBONE_ERR_ARG, "create: bad box");
    if (int rc = ensure_init()) return rc;
    auto P = std::make_unique<nekbone_problem>();
    P->nelx = nelx;
    P->nely = nely;
    P->nelz = nelz;
    P->nelt = nelt;
    P->nx1 = nx1;
    P->axv = tuned_or_default(nx1);
    P->n = int64_t(nelt) * nx1 * nx1 * nx1;
    P->lglel.assign(lglel, lglel + nelt);
    std::memcpy(P->dxm1, dxm1, sizeof(double) * nx1 * nx1);
    double a[kMaxNx * kMaxNx], cc[kMaxNx * kMaxNx], d[kMaxNx * kMaxNx];
    semhat(a, P->wgll, cc, d, P->zgll, nx1 - 1);
    P->glo.assign(glo_num, glo_num + P->n);
    mask_lists(nx1, nelt, nelx, nely, nelz, P->lglel, P->mfull, P->mref);
    if (int rc = finish_problem(P.get(), c)) {
        nekbone_problem_destroy(P.release());
        return rc;
    }
    if (gxyz) {
        if (int rc = nekbone_set_geometry(P.get(), gxyz)) {
            nekbone_problem_destroy(P.release());
            return rc;
        }
    } else if (int rc = launch_brick_g(P.get())) {
        nekbone_problem_destroy(P.release());
        return rc;
    }
    NB_HIP(hipStreamSynchronize(stream()));
    *out = P.release();
    return NEKBONE_OK;
}

extern "C" int nekbone_problem_destroy(nekbone_problem *P)
{
    if (!P) return NEKBONE_OK;
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    drop_problem_graph(P);
    dfree(P->d_D);
    dfree(P->d_Drow);
    dfree(P->d_g);
    dfree(P->d_c);
    dfree(P->d_dinv);
    dfree(P->d_flag);
    dfree(P->d_gx);
    dfree(P->d_gxbad);
    dfree(P->d_pring);
    dfree(P->d_perm);
    dfree(P->d_iref);
    dfree(P->gs.d_m2);
    dfree(P->gs.d_m4);
    dfree(P->gs.d_m8);
    dfree(P->gs.d_roff);
    dfree(P->gs.d_ridx);
    for (double **v : {&P->d_x, &P->d_r, &P->d_p, &P->d_w, &P->d_f, &P->d_t, &P->d_t2, &P->d_part,
                       &P->d_hist})
        dfree(*v);
    dfree(P->d_scal);
    free_mixed(P);
    hsmg_free(P);
    (void)timing_resolve(P);
    for (hipEvent_t e : P->timing.pool) (void)hipEventDestroy(e);
    if (P->timing.cur) (void)hipEventDestroy(P->timing.cur);
    delete P;
    return NEKBONE_OK;
}

extern "C" int64_t nekbone_problem_npts(const nekbone_problem *P) { return P ? P->n : -1; }
extern "C" int nekbone_problem_nx1(const nekbone_problem *P) { return P ? P->nx1 : -1; }
extern "C" int nekbone_problem_nelt(const nekbone_problem *P) { return P ? P->nelt : -1; }

extern "C" int nekbone_set_geometry(nekbone_problem *P, const double *gxyz)
{
    if (int rc = check(P)) return rc;
    if (!gxyz) return fail(NEKBONE_ERR_ARG, "set_geometry: null g");
    double *tmp = nullptr;
    if (int rc = dalloc(&tmp, 6 * P->n)) return rc;
    if (int rc = h2d(tmp, gxyz, sizeof(double) * 6 * size_t(P->n))) {
        dfree(tmp);
        return rc;
    }
    int rc = launch_aos_to_soa(P, tmp, P->d_g, 0);
    (void)hipStreamSynchronize(stream());
    dfree(tmp);
    P->jacobi = false;
    P->geom_version++;
    return rc;
}

extern "C" int nekbone_get_geometry(nekbone_problem *P, double *gxyz)
{
    if (int rc = check(P)) return rc;
    double *tmp = nullptr;
    if (int rc = dalloc(&tmp, 6 * P->n)) return rc;
    int rc = launch_aos_to_soa(P, tmp, P->d_g, 1);
    if (rc == NEKBONE_OK) {
        rc = d2h(gxyz, tmp, sizeof(double) * 6 * size_t(P->n));
    }
    dfree(tmp);
    return rc;
}

extern "C" int nekbone_set_geometry_deformed(nekbone_problem *P, double amp, uint64_t seed)
{
    if (int rc = check(P)) return rc;
    if (!(amp >= 0.0) || amp >= 0.25)
        return fail(NEKBONE_ERR_ARG, "deformed geometry: need 0 <= amp < 0.25");
    const int vx = P->nelx + 1, vy = P->nely + 1, vz = P->nelz + 1;
    const int64_t nv = int64_t(vx) * vy * vz;
    double *d_v = nullptr;
    if (int rc = dalloc(&d_v, 3 * nv)) return rc;
    // Vertex k of the width-2 grid, coordinate d: 2*index + (2*amp)*U_{3k+d}.
    std::vector<double> V(size_t(3 * nv));
    if (int rc = launch_rand(d_v, 3 * nv, seed)) return rc;
    NB_HIP(hipMemcpyAsync(V.data(), d_v, sizeof(double) * V.size(), hipMemcpyDeviceToHost,
                          stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    const double scale = amp * 2.0;
    for (int c = 0; c < vz; ++c)
        for (int b = 0; b < vy; ++b)
            for (int a = 0; a < vx; ++a) {
                const int64_t v = a + int64_t(vx) * (b + int64_t(vy) * c);
                V[size_t(3 * v + 0)] = 2.0 * a + scale * V[size_t(3 * v + 0)];
                V[size_t(3 * v + 1)] = 2.0 * b + scale * V[size_t(3 * v + 1)];
                V[size_t(3 * v + 2)] = 2.0 * c + scale * V[size_t(3 * v + 2)];
            }
    if (int rc = h2d(d_v, V.data(), sizeof(double) * V.size())) {
        dfree(d_v);
        return rc;
    }
    int rc = launch_deformed_g(P, d_v);
    (void)hipStreamSynchronize(stream());
    dfree(d_v);
    P->jacobi = false;
    P->geom_version++;
    return rc;
}

extern "C" int nekbone_get_dxm1(const nekbone_problem *P, double *dxm1)
{
    if (int rc = check(P)) return rc;
    std::memcpy(dxm1, P->dxm1, sizeof(double) * P->nx1 * P->nx1);
    return NEKBONE_OK;
}

extern "C" int nekbone_get_c(nekbone_problem *P, double *c)
{
    if (int rc = check(P)) return rc;
    return download_ref(P, c, P->d_c);
}

extern "C" int nekbone_get_glo(const nekbone_problem *P, int64_t *glo)
{
    if (int rc = check(P)) return rc;
    std::memcpy(glo, P->glo.data(), sizeof(int64_t) * size_t(P->n));
    return NEKBONE_OK;
}

extern "C" int64_t nekbone_gs_ngroups(const nekbone_problem *P) { return P ? P->gs.ngroups : -1; }
extern "C" int64_t nekbone_gs_nshared(const nekbone_problem *P) { return P ? P->gs.nshared : -1; }

extern "C" int nekbone_set_mask_mode(nekbone_problem *P, int mode)
{
    if (int rc = check(P)) return rc;
    if (mode != NEKBONE_MASK_REF && mode != NEKBONE_MASK_DIRICHLET)
        return fail(NEKBONE_ERR_ARG, "bad mask mode");
    P->mask_mode = mode;
    return NEKBONE_OK;
}

extern "C" int nekbone_reset_mask(nekbone_problem *P)
{
    if (int rc = check(P)) return rc;
    P->mask_built = 0;
    return NEKBONE_OK;
}

extern "C" int nekbone_set_arith(nekbone_problem *P, int arith)
{
    if (int rc = check(P)) return rc;
    if (arith != NEKBONE_ARITH_EXACT && arith != NEKBONE_ARITH_FMA)
        return fail(NEKBONE_ERR_ARG, "bad arithmetic mode");
    P->arith = arith;
    return NEKBONE_OK;
}

extern "C" int nekbone_get_variant(const nekbone_problem *P, int *epb, int *pf, int *xu)
{
    if (int rc = check(P)) return rc;
    if (epb) *epb = P->axv.epb;
    if (pf) *pf = P->axv.pf;
    if (xu) *xu = P->axv.xu;
    return NEKBONE_OK;
}

extern "C" int nekbone_set_variant(nekbone_problem *P, int epb, int pf, int xu)
{
    if (int rc = check(P)) return rc;
    for (const AxVariant &v : ax_variants(P->nx1))
        if (v.epb == epb && v.pf == pf && v.xu == xu) {
            P->axv = v;
            return NEKBONE_OK;
        }
    return fail(NEKBONE_ERR_ARG, "set_variant: no such kernel shape for this order");
}

int nb::tune_rhs(nekbone_problem *P, uint64_t seed)
{
    if (int rc = launch_rand(P->d_f, P->n, seed)) return rc;
    return launch_gs(P, P->d_f, nullptr);
}

extern "C" int nekbone_autotune(nekbone_problem *P, int precond, double *best_ms)
{
    if (int rc = check(P)) return rc;
    if (precond == NEKBONE_PRECOND_JACOBI && !P->jacobi)
        return fail(NEKBONE_ERR_STATE, "autotune: Jacobi not set up");
    if (precond != NEKBONE_PRECOND_JACOBI && precond != NEKBONE_PRECOND_IDENTITY)
        return fail(NEKBONE_ERR_ARG, "autotune: identity or Jacobi");
    std::vector<AxVariant> cands = ax_variants(P->nx1);
    if (cands.empty()) cands.push_back(ax_default_variant(P->nx1));
    const AxVariant keep = P->axv;
    const int mbit = P->mask_mode == NEKBONE_MASK_DIRICHLET ? 1 : 2;
    const int warm = 2, reps = 6;
    if (int rc = ensure_hist(P, warm + reps)) return rc;
    if (int rc = tune_rhs(P, 12345)) return rc;
    hipEvent_t e0, e1;
    NB_HIP(hipEventCreate(&e0));
    NB_HIP(hipEventCreate(&e1));
    double best = 1e300;
    AxVariant win = keep;
    int rc = NEKBONE_OK;
    for (const AxVariant &v : cands) {
        P->axv = v;
        rc = launch_cg_init(P, P->d_x, P->d_r, P->d_f, 1, precond, 0.0, cg_zform(P, precond));
        int k = 0;
        for (int i = 0; i < warm && !rc; ++i)
            rc = launch_cg_iteration(P, P->d_x, P->d_r, precond, mbit, k++);
        if (!rc && hipEventRecord(e0, stream()) != hipSuccess) rc = NEKBONE_ERR_HIP;
        for (int i = 0; i < reps && !rc; ++i)
            rc = launch_cg_iteration(P, P->d_x, P->d_r, precond, mbit, k++);
        if (!rc && hipEventRecord(e1, stream()) != hipSuccess) rc = NEKBONE_ERR_HIP;
        if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = NEKBONE_ERR_HIP;
        float ms = 0.f;
        if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = NEKBONE_ERR_HIP;
        if (rc) break;
        if (ms / reps < best) {
            best = ms / reps;
            win = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        P->axv = keep;
        return rc == NEKBONE_ERR_HIP ? fail(rc, "autotune: HIP error") : rc;
    }
    P->axv = win;
    tuned_cache()[std::make_tuple(P->nx1, precond, P->arith)] = win;
    tuned_cache()[std::make_tuple(P->nx1, -1, -1)] = win;
    if (best_ms) *best_ms = best;
    return NEKBONE_OK;
}

// Tuning cache on disk: the element-kernel shapes nekbone_autotune found,
// as JSON, one entry per (order, preconditioner, arithmetic); -1/-1 is the
// shape later problems of that order start with.
extern "C" int nekbone_tuning_save(const char *path)
{
    if (!path) return fail(NEKBONE_ERR_ARG, "tuning_save: null path");
    FILE *fp = std::fopen(path, "w");
    if (!fp) return fail(NEKBONE_ERR_ARG, std::string("tuning_save: cannot write ") + path);
    std::fprintf(fp, "{\"format\": \"nekbone_amd tuning 1\", \"element_kernel\": [");
    bool first = true;
    for (const auto &kv : tuned_cache()) {
        int nx1, precond, arith;
        std::tie(nx1, precond, arith) = kv.first;
        std::fprintf(fp, "%s\n  {\"nx1\": %d, \"precond\": %d, \"arith\": %d, \"epb\": %d, "
                         "\"pf\": %d, \"xu\": %d}",
                     first ? "" : ",", nx1, precond, arith, kv.second.epb, kv.second.pf,
                     kv.second.xu);
        first = false;
    }
    std::fprintf(fp, "\n]}\n");
    const bool ok = std::fclose(fp) == 0;
    return ok ? NEKBONE_OK : fail(NEKBONE_ERR_ARG, "tuning_save: write failed");
}

// Reads what nekbone_tuning_save wrote; shapes not compiled for their order
// are skipped.  *loaded: the number of entries taken.
extern "C" int nekbone_tuning_load(const char *path, int *loaded)
{
    if (!path) return fail(NEKBONE_ERR_ARG, "tuning_load: null path");
    FILE *fp = std::fopen(path, "r");
    if (!fp) return fail(NEKBONE_ERR_ARG, std::string("tuning_load: cannot read ") + path);
    std::string text;
    char buf[4096];
    for (size_t got; (got = std::fread(buf, 1, sizeof buf, fp)) > 0;) text.append(buf, got);
    std::fclose(fp);
    if (text.find("\"nekbone_amd tuning 1\"") == std::string::npos)
        return fail(NEKBONE_ERR_ARG, "tuning_load: not a tuning file");
    int count = 0;
    for (size_t at = text.find('{', 1); at != std::string::npos; at = text.find('{', at + 1)) {
        int nx1, precond, arith, epb, pf, xu;
        if (std::sscanf(text.c_str() + at,
                        "{\"nx1\": %d, \"precond\": %d, \"arith\": %d, \"epb\": %d, \"pf\": %d, "
                        "\"xu\": %d}",
                        &nx1, &precond, &arith, &epb, &pf, &xu) != 6)
            continue;
        bool compiled = false;
        for (const AxVariant &v : ax_variants(nx1))
            compiled = compiled || (v.epb == epb && v.pf == pf && v.xu == xu);
        if (!compiled) continue;
        tuned_cache()[std::make_tuple(nx1, precond, arith)] = AxVariant{epb, pf, xu};
        ++count;
    }
    if (loaded) *loaded = count;
    return NEKBONE_OK;
}

extern "C" int nekbone_set_graph(nekbone_problem *P, int on)
{
    if (int rc = check(P)) return rc;
    P->use_graph = on != 0;
    if (!P->use_graph) drop_problem_graph(P);
    return NEKBONE_OK;
}

extern "C" int nekbone_set_fused(nekbone_problem *P, int on)
{
    if (int rc = check(P)) return rc;
    P->fused = on != 0;
    return NEKBONE_OK;
}

extern "C" int nekbone_set_xring(nekbone_problem *P, int on)
{
    if (int rc = check(P)) return rc;
    if (on != 0 && on != 1) return fail(NEKBONE_ERR_ARG, "set_xring: on must be 0 or 1");
    drop_problem_graph(P);
    P->xring = on;
    if (!on) {
        dfree(P->d_pring);
        P->ring_on = false;
    }
    return NEKBONE_OK;
}

extern "C" int nekbone_get_cg_form(nekbone_problem *P, int *gather, int *xring)
{
    if (int rc = check(P)) return rc;
    if (gather) *gather = cg_gather(P) ? 1 : 0;
    if (xring) *xring = P->ring_on ? 1 : 0;
    return NEKBONE_OK;
}

extern "C" int nekbone_set_gather(nekbone_problem *P, int on, int *active)
{
    if (int rc = check(P)) return rc;
    if (on == 0 || on == 1) P->gather = on;
    else if (on != -1) return fail(NEKBONE_ERR_ARG, "set_gather: on must be 0, 1 or -1");
    if (active) *active = cg_gather(P) ? 1 : 0;
    return NEKBONE_OK;
}

extern "C" int nekbone_setup_jacobi(nekbone_problem *P)
{
    if (int rc = check(P)) return rc;
    if (!P->d_dinv)
        if (int rc = dalloc(&P->d_dinv, P->n)) return rc;
    if (int rc = launch_jacobi_local(P, P->d_dinv)) return rc;
    if (int rc = launch_gs(P, P->d_dinv, nullptr)) return rc;
    if (int rc = launch_jacobi_finish(P, P->d_dinv)) return rc;
    NB_HIP(hipStreamSynchronize(stream()));
    P->jacobi = true;
    P->dinv_version++;
    return NEKBONE_OK;
}

extern "C" int nekbone_get_dinv(nekbone_problem *P, double *dinv)
{
    if (int rc = check(P)) return rc;
    if (!P->d_dinv) return fail(NEKBONE_ERR_STATE, "Jacobi not set up");
    return download_ref(P, dinv, P->d_dinv);
}

// ======================================================================
// Operators.  Caller vectors are in reference order; the work runs on
// device-order copies (d_t2 in, d_w out).
// ======================================================================
static int ax_dev(nekbone_problem *P)
{
    if (int rc = launch_ax_local(P, P->d_w, P->d_t2)) return rc;
    if (int rc = launch_gs(P, P->d_w, nullptr)) return rc;
    return launch_post(P, P->d_w, P->d_t2, next_mask_bit(P), false);
}

extern "C" int nekbone_ax_d(nekbone_problem *P, double *w, const double *u)
{
    if (int rc = check(P)) return rc;
    if (int rc = launch_to_dev(P, P->d_t2, u)) return rc;
    if (int rc = ax_dev(P)) return rc;
    return launch_to_ref(P, w, P->d_w);
}

extern "C" int nekbone_ax(nekbone_problem *P, double *w, const double *u)
{
    if (int rc = check(P)) return rc;
    if (int rc = upload_dev(P, P->d_t2, u)) return rc;
    if (int rc = ax_dev(P)) return rc;
    return download_ref(P, w, P->d_w);
}

extern "C" int nekbone_ax_local(nekbone_problem *P, double *w, const double *u)
{
    if (int rc = check(P)) return rc;
    if (int rc = upload_dev(P, P->d_t2, u)) return rc;
    if (int rc = launch_ax_local(P, P->d_w, P->d_t2)) return rc;
    return download_ref(P, w, P->d_w);
}

extern "C" int nekbone_dssum_d(nekbone_problem *P, double *u)
{
    if (int rc = check(P)) return rc;
    if (int rc = launch_to_dev(P, P->d_t2, u)) return rc;
    if (int rc = launch_gs(P, P->d_t2, nullptr)) return rc;
    return launch_to_ref(P, u, P->d_t2);
}

extern "C" int nekbone_dssum(nekbone_problem *P, double *u)
{
    if (int rc = check(P)) return rc;
    if (int rc = upload_dev(P, P->d_t2, u)) return rc;
    if (int rc = launch_gs(P, P->d_t2, nullptr)) return rc;
    return download_ref(P, u, P->d_t2);
}

extern "C" int nekbone_maskit(nekbone_problem *P, double *w)
{
    if (int rc = check(P)) return rc;
    if (int rc = upload_dev(P, P->d_t2, w)) return rc;
    if (int rc = launch_mask(P, P->d_t2, next_mask_bit(P))) return rc;
    return download_ref(P, w, P->d_t2);
}

extern "C" int nekbone_glsc3(nekbone_problem *P, const double *a, const double *b,
                             const double *m, double *result)
{
    if (int rc = check(P)) return rc;
    const size_t bytes = sizeof(double) * size_t(P->n);
    NB_HIP(hipMemcpyAsync(P->d_t, a, bytes, hipMemcpyHostToDevice, stream()));
    NB_HIP(hipMemcpyAsync(P->d_w, b, bytes, hipMemcpyHostToDevice, stream()));
    NB_HIP(hipMemcpyAsync(P->d_f, m, bytes, hipMemcpyHostToDevice, stream()));
    if (int rc = ensure_hist(P, 0)) return rc;
    if (int rc = launch_glsc3(P, P->d_t, P->d_w, P->d_f, P->d_hist)) return rc;
    NB_HIP(hipMemcpyAsync(result, P->d_hist, sizeof(double), hipMemcpyDeviceToHost, stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

// ======================================================================
// CG
// ======================================================================
static int copy_hist(nekbone_problem *P, double *hist, int niter)
{
    if (!hist) return NEKBONE_OK;
    NB_HIP(hipMemcpyAsync(hist, P->d_hist, sizeof(double) * size_t(1 + 3 * niter),
                          hipMemcpyDeviceToHost, stream()));
    return NEKBONE_OK;
}

static int check_status(nekbone_problem *P)
{
    CgScalars s;
    NB_HIP(hipMemcpyAsync(&s, P->d_scal, sizeof s, hipMemcpyDeviceToHost, stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    if (s.status) return fail(NEKBONE_ERR_NUMERIC, "cg: non-positive or non-finite pap");
    return NEKBONE_OK;
}

extern "C" int nekbone_cg_d(nekbone_problem *P, double *x, const double *f, double *r,
                            int niter, int precond, double *hist)
{
    if (int rc = check(P)) return rc;
    if (int rc = launch_to_dev(P, P->d_f, f)) return rc;
    if (int rc = run_cg(P, P->d_x, P->d_f, P->d_r, niter, precond, 0.0)) return rc;
    if (int rc = copy_hist(P, hist, niter)) return rc;
    if (x)
        if (int rc = launch_to_ref(P, x, P->d_x)) return rc;
    if (r)
        if (int rc = launch_to_ref(P, r, P->d_r)) return rc;
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

extern "C" int nekbone_cg(nekbone_problem *P, double *x, const double *f, double *r, int niter,
                          int precond, double *hist)
{
    if (int rc = check(P)) return rc;
    if (int rc = upload_dev(P, P->d_f, f)) return rc;
    if (int rc = run_cg(P, P->d_x, P->d_f, P->d_r, niter, precond, 0.0)) return rc;
    if (int rc = copy_hist(P, hist, niter)) return rc;
    if (x)
        if (int rc = download_ref(P, x, P->d_x)) return rc;
    if (r)
        if (int rc = download_ref(P, r, P->d_r)) return rc;
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

extern "C" int nekbone_cg_tol_d(nekbone_problem *P, double *x, const double *f, double *r,
                                double tol, int maxiter, int precond, int *iters,
                                double *final_rtr)
{
    if (int rc = check(P)) return rc;
    if (!(tol > 0.0)) return fail(NEKBONE_ERR_ARG, "cg_tol: tol must be > 0");
    if (int rc = check_precond(P, precond)) return rc;
    if (int rc = ensure_hist(P, maxiter)) return rc;
    if (int rc = launch_to_dev(P, P->d_f, f)) return rc;
    const int mbit0 = next_mask_bit(P);
    if (int rc = launch_cg_init(P, P->d_x, P->d_r, P->d_f, mbit0, precond, tol * tol,
                                cg_zform(P, precond)))
        return rc;
    const int mbit = next_mask_bit(P);
    // Iterations are queued in chunks, each followed by a copy of the
    // scalars into pinned memory; the next chunk is queued before the host
    // waits for the previous one's copy, so the GPU never waits for the
    // host check.  The kernels of a converged solve return at once, so the
    // chunk queued past convergence costs only its launches; with the HSMG
    // preconditioner those are some 25 per iteration, more than the host
    // check costs, so there the host checks each chunk before queueing the
    // next.
    const int chunk = 10;
    const bool ahead = precond != NEKBONE_PRECOND_HSMG;
    CgScalars s{};
    int queued = 0, posted = 0, seen = 0;
    auto queue = [&]() -> int {
        const int todo = std::min(chunk, maxiter - queued);
        if (int rc = run_iterations(P, P->d_x, P->d_r, precond, mbit, todo, queued)) return rc;
        queued += todo;
        return scal_post(P, posted++ & 1);
    };
    if (maxiter > 0) {
        if (int rc = queue()) return rc;
        for (;;) {
            if (ahead && queued < maxiter)
                if (int rc = queue()) return rc;
            if (int rc = scal_wait(seen++ & 1, s)) return rc;
            if (s.done || s.status) break;
            if (seen == posted) {
                if (queued >= maxiter) break;
                if (int rc = queue()) return rc;
            }
        }
    }
    if (int rc = launch_cg_finish(P, P->d_x, P->d_r, cg_zform(P, precond))) return rc;
    if (x)
        if (int rc = launch_to_ref(P, x, P->d_x)) return rc;
    if (r)
        if (int rc = launch_to_ref(P, r, P->d_r)) return rc;
    NB_HIP(hipMemcpyAsync(&s, P->d_scal, sizeof s, hipMemcpyDeviceToHost, stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    if (iters) *iters = s.iter;
    if (final_rtr) *final_rtr = s.rtr;
    if (s.status) return fail(NEKBONE_ERR_NUMERIC, "cg: non-positive or non-finite pap");
    if (!s.done)
        return fail(NEKBONE_NOT_CONVERGED, "cg_tol: stopped at maxiter with rtr > tol^2 rtr0");
    return NEKBONE_OK;
}

// ======================================================================
// Device buffers and synthetic inputs
// ======================================================================
extern "C" int nekbone_alloc_vec(nekbone_problem *P, double **dptr)
{
    if (int rc = check(P)) return rc;
    if (int rc = dalloc(dptr, P->n)) return rc;
    if (int rc = dzero(*dptr, sizeof(double) * size_t(P->n))) {
        dfree(*dptr);
        return rc;
    }
    return NEKBONE_OK;
}

extern "C" int nekbone_free_vec(nekbone_problem *P, double *dptr)
{
    (void)P;
    dfree(dptr);
    return NEKBONE_OK;
}

extern "C" int nekbone_h2d(nekbone_problem *P, double *dst, const double *src, int64_t count)
{
    (void)P;
    if (int rc = ensure_init()) return rc;
    NB_HIP(hipMemcpyAsync(dst, src, sizeof(double) * size_t(count), hipMemcpyHostToDevice,
                          stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

extern "C" int nekbone_d2h(nekbone_problem *P, double *dst, const double *src, int64_t count)
{
    (void)P;
    if (int rc = ensure_init()) return rc;
    NB_HIP(hipMemcpyAsync(dst, src, sizeof(double) * size_t(count), hipMemcpyDeviceToHost,
                          stream()));
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

extern "C" int nekbone_rand_uniform_d(nekbone_problem *P, double *dst, int64_t count,
                                      uint64_t seed)
{
    (void)P;
    if (int rc = ensure_init()) return rc;
    return launch_rand(dst, count, seed);
}

extern "C" int nekbone_rhs_from_raw_d(nekbone_problem *P, double *f)
{
    if (int rc = check(P)) return rc;
    if (int rc = launch_to_dev(P, P->d_t2, f)) return rc;
    if (int rc = launch_gs(P, P->d_t2, nullptr)) return rc;
    if (int rc = launch_col2(P, P->d_t2, P->d_c)) return rc;
    return launch_to_ref(P, f, P->d_t2);
}

// set_f (src/driver.f:88-101): f(i) = sin(1e9*cos(1e9*(i*i))), i = 1..n,
// where i*i is the reference's default-INTEGER product: it overflows for
// i > 46340 and the flang-built reference wraps it (two's complement), as
// the compiled reference's own output confirms.  sin and cos are the C
// library's (the arguments reach 1e18, so the result depends on the
// library's argument reduction; computed on the host, never on the
// device).  Independent points, split over host threads.
extern "C" int nekbone_set_f_raw(int64_t n, double *raw)
{
    if (n < 0 || (n > 0 && !raw)) return fail(NEKBONE_ERR_ARG, "set_f_raw: bad arguments");
    if (n > (int64_t(1) << 31) - 1) return fail(NEKBONE_ERR_ARG, "set_f_raw: n >= 2^31");
    const int nt = int(std::max<int64_t>(1, std::min<int64_t>(16, n / 65536)));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([=] {
            const int64_t a = n * t / nt, b = n * (t + 1) / nt;
            for (int64_t l = a; l < b; ++l) {
                const uint32_t i = uint32_t(l + 1);
                const int32_t ii = int32_t(i * i);  // wraps like the reference's INTEGER product
                raw[l] = std::sin(1.e9 * std::cos(1.e9 * double(ii)));
            }
        });
    for (auto &x : th) x.join();
    return NEKBONE_OK;
}

extern "C" int nekbone_set_f_d(nekbone_problem *P, double *f)
{
    if (int rc = check(P)) return rc;
    if (!f) return fail(NEKBONE_ERR_ARG, "set_f: null vector");
    std::vector<double> raw(static_cast<size_t>(P->n));
    if (int rc = nekbone_set_f_raw(P->n, raw.data())) return rc;
    if (int rc = h2d(f, raw.data(), sizeof(double) * raw.size())) return rc;
    return nekbone_rhs_from_raw_d(P, f);  // dssum, then col2 with c (:97-98)
}

// ======================================================================
// Timing
// ======================================================================
extern "C" int nekbone_timing_enable(nekbone_problem *P, int mask)
{
    if (int rc = check(P)) return rc;
    if (int rc = timing_resolve(P)) return rc;
    P->timing.mask = mask & 31;
    return NEKBONE_OK;
}

extern "C" int nekbone_timing_get(nekbone_problem *P, int kclass, int64_t *launches, double *ms)
{
    if (int rc = check(P)) return rc;
    if (kclass < 0 || kclass > 4) return fail(NEKBONE_ERR_ARG, "bad kernel class");
    if (int rc = timing_resolve(P)) return rc;
    if (launches) *launches = P->timing.launches[kclass];
    if (ms) *ms = P->timing.ms[kclass];
    return NEKBONE_OK;
}

namespace nb {
namespace {
__global__ void k_copy2(double2 *__restrict__ dst, const double2 *__restrict__ src, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}
}  // namespace
}  // namespace nb

extern "C" int nekbone_bw_probe(int64_t bytes, int reps, double *ms)
{
    if (int rc = ensure_init()) return rc;
    if (bytes < 1024 || reps < 1) return fail(NEKBONE_ERR_ARG, "bw_probe: bytes >= 1024, reps >= 1");
    const int64_t n2 = bytes / 2 / 16;  // double2 elements per buffer
    double2 *a = nullptr, *b = nullptr;
    NB_HIP(hipMalloc(&a, sizeof(double2) * size_t(n2)));
    if (hipMalloc(&b, sizeof(double2) * size_t(n2)) != hipSuccess) {
        (void)hipFree(a);
        return fail(NEKBONE_ERR_HIP, "bw_probe: allocation");
    }
    (void)hipMemsetAsync(a, 0, sizeof(double2) * size_t(n2), stream());
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const unsigned blocks = unsigned((n2 + 255) / 256);
    hipLaunchKernelGGL(k_copy2, dim3(blocks), dim3(256), 0, stream(), b, a, n2);
    (void)hipEventRecord(e0, stream());
    for (int i = 0; i < reps; ++i)
        hipLaunchKernelGGL(k_copy2, dim3(blocks), dim3(256), 0, stream(), (i & 1) ? a : b,
                           (i & 1) ? b : a, n2);
    (void)hipEventRecord(e1, stream());
    hipError_t err = hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return fail(NEKBONE_ERR_HIP, "bw_probe: kernel");
    if (ms) *ms = double(t) / reps;
    return NEKBONE_OK;
}

extern "C" int nekbone_timing_reset(nekbone_problem *P)
{
    if (int rc = check(P)) return rc;
    if (int rc = timing_resolve(P)) return rc;
    for (int k = 0; k < 5; ++k) {
        P->timing.launches[k] = 0;
        P->timing.ms[k] = 0.0;
    }
    return NEKBONE_OK;
}
