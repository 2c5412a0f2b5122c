// HSMG preconditioner application on the device: h1mg_solve
// (src/hsmg.f:960-1034), an additive two-level Schwarz/FDM V-cycle with a
// trivial coarse solve (QQ^T on the order-1 level).
//
// Per Schwarz level (h1mg_schwarz, :1036-1095, sigma = 1):
//   k_fdm     mask r in place (the reference's side effect, :1067), build
//             the extended element (own points, ghost face = the
//             neighbour's point one inside, as the (n+2)^3 exchange of
//             :1071-1076 delivers it), local solve e = S D^-1 S^T r with
//             S = S_z x S_y x S_x (h1mg_do_fast, :1297-1320); writes the
//             own part and the ghost faces of e
//   k_swz_ovl add the neighbours' ghost results onto the first interior
//   k_swz_gs  layer (:1081-1085), dssum on the level (:1089), mask, then
//             the square-rooted Schwarz weights (:1090, :1144-1178): one
//             pass over the points, one over the level's gs groups
// Restriction k_rstr (h1mg_rstr, :1340-1357) and interpolation-add
// k_intp (h1mg_intp, :1436-1443) are per-element tensor products through
// LDS.  Every sum keeps the reference's order (mxm: left to right from the
// first product; the ghost exchange as (own + neighbour) - own).
#include <vector>

#include "nb_device.h"

namespace nb {
namespace {

constexpr int kT = 256;
// threads per element of the restriction and interpolation kernels
constexpr int kTE = 128;

struct FdmArgs {
    double *r;
    const uint8_t *bnd;
    const int32_t *pos, *gsrc, *gin;
    const int16_t *eop;
    double *eown, *ghost;
    int nelt;
    const CgScalars *guard;  // may be null: return at once when guard->done
};

// out(i,jk) = sum_m M(i,m) in(m,jk)       (first index)
__device__ void tx(double *out, const double *in, const double *M, int nv, int nu, int n23)
{
    for (int q = threadIdx.x; q < nv * n23; q += blockDim.x) {
        const int i = q % nv, jk = q / nv;
        double s = M[i] * in[nu * jk];
        for (int m = 1; m < nu; ++m) s = s + M[i + nv * m] * in[m + nu * jk];
        out[q] = s;
    }
}
// out(i,j,k) = sum_m in(i,m,k) Bt(m,j)     (second index); in is (n1, nu, n3)
__device__ void ty(double *out, const double *in, const double *Bt, int n1, int nv, int nu, int n3)
{
    for (int q = threadIdx.x; q < n1 * nv * n3; q += blockDim.x) {
        const int i = q % n1, j = (q / n1) % nv, k = q / (n1 * nv);
        const double *row = in + n1 * nu * k + i;
        double s = row[0] * Bt[nu * j];
        for (int m = 1; m < nu; ++m) s = s + row[n1 * m] * Bt[m + nu * j];
        out[q] = s;
    }
}
// out(ij,k) = sum_m in(ij,m) Ct(m,k)       (third index); in is (n12, nu)
__device__ void tz(double *out, const double *in, const double *Ct, int n12, int nv, int nu)
{
    for (int q = threadIdx.x; q < n12 * nv; q += blockDim.x) {
        const int ij = q % n12, k = q / n12;
        double s = in[ij] * Ct[nu * k];
        for (int m = 1; m < nu; ++m) s = s + in[ij + n12 * m] * Ct[m + nu * k];
        out[q] = s;
    }
}

// Sizes of the local FDM solve on a level with NH points per direction:
// the extended element has NL = NH+2 points per direction; one thread per
// grid line (NL^2 per element), EPB elements per workgroup.
template <int NH>
struct FdmShape {
    static constexpr int NL = NH + 2, L2 = NL * NL, L3 = L2 * NL;
    // block layout in LDS: point (i, j, k) at i + P1 j + P2 k, the j stride
    // odd so that the lanes of an x-line pass (one line (j, k) each) fall in
    // distinct banks
    static constexpr int P1 = NL | 1, P2 = P1 * NL, LB = P2 * NL;
    static constexpr int EPB = L2 >= 256 ? 1 : 256 / L2;
    static constexpr int NT = (EPB * L2 + 63) / 64 * 64;
    static constexpr int OPS = 2 * L2 + NL;  // S, S^T (column-major), eigenvalues
};

// v <- C v for a column-major NL x NL matrix C in LDS: out_i is summed
// left to right from the first product, C(i,0) v_0 + C(i,1) v_1 + ...
// (the mxm order), each step one fused multiply-add.
template <typename T, int NL>
__device__ __forceinline__ void line_apply(T (&v)[NL], const T *__restrict__ C)
{
    T o[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) o[i] = C[i] * v[0];
#pragma unroll
    for (int m = 1; m < NL; ++m)
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            if constexpr (sizeof(T) == 4)
                o[i] = __builtin_fmaf(C[i + NL * m], v[m], o[i]);
            else
                o[i] = __builtin_fma(C[i + NL * m], v[m], o[i]);
        }
#pragma unroll
    for (int i = 0; i < NL; ++i) v[i] = o[i];
}

template <typename T, int NL>
__device__ __forceinline__ void line_load(T (&v)[NL], const T *b, int base, int stride)
{
#pragma unroll
    for (int m = 0; m < NL; ++m) v[m] = b[base + stride * m];
}

template <typename T, int NL>
__device__ __forceinline__ void line_store(const T (&v)[NL], T *b, int base, int stride)
{
#pragma unroll
    for (int m = 0; m < NL; ++m) b[base + stride * m] = v[m];
}

// Extended-block position (in the LDS layout of FdmShape) of ghost slot t
// of face f (x-, x+, y-, y+, z-, z+): the face's (a, b) = (t % NH + 1,
// t / NH + 1) in its two tangential directions, the normal coordinate 0 or
// NL-1.
template <int NH>
__device__ __forceinline__ int ghost_ext(int f, int t)
{
    constexpr int NL = NH + 2, P1 = NL | 1, P2 = P1 * NL;
    const int a = t % NH + 1, b = t / NH + 1, side = (f & 1) ? NL - 1 : 0;
    if (f < 2) return side + P1 * a + P2 * b;
    if (f < 4) return a + P1 * side + P2 * b;
    return a + P1 * b + P2 * side;
}

// Extended-block position of own point p = i + NH j + NH^2 k.
template <int NH>
__device__ __forceinline__ int own_ext(int p)
{
    constexpr int NL = NH + 2, P1 = NL | 1, P2 = P1 * NL;
    const int i = p % NH, j = (p / NH) % NH, k = p / (NH * NH);
    return (i + 1) + P1 * (j + 1) + P2 * (k + 1);
}

// Local solves of one Schwarz level (h1mg_schwarz_part1, src/hsmg.f:1052-1079):
//   mask r in place (the reference's side effect, :1067); build the
//   extended element (own points, each ghost face = the neighbour's point
//   one inside, delivered as (own + neighbour) - own like the (n+2)^3
//   exchange of :1071-1076; edges and corners zero); e = S D^-1 S^T r with
//   S = S_z x S_y x S_x (h1mg_do_fast, :1297-1320): contractions along x,
//   y, z with the S^T factors, the diagonal (l_x + l_y + l_z, cut to zero
//   at or below the element's eps), then x, y, z with the S factors; write
//   the own part and the ghost faces of e.
// The element's block sits in LDS; each contraction reads a thread's line
// into registers and writes it back in place.  T is the arithmetic of the
// local solve (FP64, or FP32 for the single-precision smoother); the level
// vectors stay FP64.
template <typename T, int NH>
__global__ void __launch_bounds__(FdmShape<NH>::NT)
    k_fdm(FdmArgs a, const T *__restrict__ ops, const T *__restrict__ epsv)
{
    if (a.guard && a.guard->done) return;
    using S = FdmShape<NH>;
    constexpr int NL = S::NL, L2 = S::L2, EPB = S::EPB, OPS = S::OPS;
    constexpr int P1 = S::P1, P2 = S::P2, L3 = S::LB;
    constexpr int H2 = NH * NH, N3 = H2 * NH, G2 = 6 * H2;
    __shared__ T blk[EPB][L3];
    __shared__ T op[EPB][3][OPS];
    const int e0 = blockIdx.x * EPB, tid = threadIdx.x, nelt = a.nelt;
    // Every global load (operators, own points, ghost sources) is issued
    // before the first barrier, so they are all in flight together; the
    // LDS is cleared meanwhile (edges and corners of the extended block
    // stay zero).  Dirichlet points read as zero; r itself is masked at
    // the end.
    constexpr int IOP = (EPB * 3 * OPS + S::NT - 1) / S::NT;
    T vop[IOP];
#pragma unroll
    for (int it = 0; it < IOP; ++it) {
        const int q = tid + it * S::NT, el = q / (3 * OPS), w = q - el * (3 * OPS), d = w / OPS;
        vop[it] = T(0);
        if (q < EPB * 3 * OPS && e0 + el < nelt)
            vop[it] = ops[int64_t(a.eop[3 * (e0 + el) + d]) * OPS + (w - d * OPS)];
    }
    constexpr int IO = (EPB * N3 + S::NT - 1) / S::NT, IG = (EPB * G2 + S::NT - 1) / S::NT;
    // own points: memory index and mask kept for the stores at the end
    double vown[IO], vgh[IG];
    int64_t mo[IO];
    uint8_t bo[IO];
#pragma unroll
    for (int it = 0; it < IO; ++it) {
        const int q = tid + it * S::NT, el = q / N3, e = e0 + el;
        mo[it] = -1;
        bo[it] = 0;
        vown[it] = 0.0;
        if (q < EPB * N3 && e < nelt) {
            mo[it] = int64_t(e) * N3 + a.pos[q - el * N3];
            bo[it] = a.bnd[mo[it]];
            vown[it] = a.r[mo[it]];
        }
    }
#pragma unroll
    for (int it = 0; it < IG; ++it) {
        const int q = tid + it * S::NT, el = q / G2, e = e0 + el;
        vgh[it] = 0.0;
        if (q < EPB * G2 && e < nelt) {
            const int64_t slot = int64_t(e) * G2 + (q - el * G2);
            const int32_t sn = a.gsrc[slot];
            if (sn >= 0) {
                const int32_t o = a.gin[slot];
                const uint8_t bs = a.bnd[sn], bi = a.bnd[o];
                const double ro = a.r[o], rs = a.r[sn];
                const double vo = bi ? 0.0 : ro;
                const double vs = bs ? 0.0 : rs;
                vgh[it] = (vo + vs) - vo;
            }
        }
    }
#pragma unroll
    for (int it = 0; it < IO; ++it)
        if (bo[it]) vown[it] = 0.0;
    for (int q = tid; q < EPB * L3; q += S::NT) (&blk[0][0])[q] = T(0);
#pragma unroll
    for (int it = 0; it < IOP; ++it) {
        const int q = tid + it * S::NT;
        if (q < EPB * 3 * OPS) (&op[0][0][0])[q] = vop[it];
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IO; ++it) {
        const int q = tid + it * S::NT, el = q / N3;
        if (q < EPB * N3 && e0 + el < nelt) blk[el][own_ext<NH>(q - el * N3)] = T(vown[it]);
    }
#pragma unroll
    for (int it = 0; it < IG; ++it) {
        const int q = tid + it * S::NT, el = q / G2;
        if (q < EPB * G2 && e0 + el < nelt) {
            const int w = q - el * G2;
            blk[el][ghost_ext<NH>(w / H2, w - (w / H2) * H2)] = T(vgh[it]);
        }
    }
    __syncthreads();
    const int el = tid / L2, l = tid - el * L2, la = l % NL, lb = l / NL;
    const bool act = el < EPB && e0 + el < nelt;
    T *b = blk[act ? el : 0];
    const T *ox = op[act ? el : 0][0], *oy = op[act ? el : 0][1], *oz = op[act ? el : 0][2];
    T v[NL];
    // S^T r: x lines (la, lb) = (j, k), y lines (i, k), z lines (i, j)
    if (act) {
        line_load(v, b, P1 * la + P2 * lb, 1);
        line_apply<T, NL>(v, ox + L2);
        line_store(v, b, P1 * la + P2 * lb, 1);
    }
    __syncthreads();
    if (act) {
        line_load(v, b, la + P2 * lb, P1);
        line_apply<T, NL>(v, oy + L2);
        line_store(v, b, la + P2 * lb, P1);
    }
    __syncthreads();
    if (act) {
        line_load(v, b, la + P1 * lb, P2);
        line_apply<T, NL>(v, oz + L2);
        const T eps = epsv[e0 + el], lxy = ox[2 * L2 + la] + oy[2 * L2 + lb];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const T diag = lxy + oz[2 * L2 + k];
            const T d = (diag > eps) ? T(1) / diag : T(0);
            v[k] = d * v[k];
        }
        line_store(v, b, la + P1 * lb, P2);
    }
    __syncthreads();
    if (act) {
        line_load(v, b, P1 * la + P2 * lb, 1);
        line_apply<T, NL>(v, ox);
        line_store(v, b, P1 * la + P2 * lb, 1);
    }
    __syncthreads();
    if (act) {
        line_load(v, b, la + P2 * lb, P1);
        line_apply<T, NL>(v, oy);
        line_store(v, b, la + P2 * lb, P1);
    }
    __syncthreads();
    if (act) {
        line_load(v, b, la + P1 * lb, P2);
        line_apply<T, NL>(v, oz);
        line_store(v, b, la + P1 * lb, P2);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IO; ++it) {
        const int q = tid + it * S::NT, el2 = q / N3;
        if (mo[it] < 0) continue;
        a.eown[mo[it]] = double(blk[el2][own_ext<NH>(q - el2 * N3)]);
        if (bo[it]) a.r[mo[it]] = 0.0;  // the caller's r leaves masked (src/hsmg.f:1067)
    }
    for (int q = tid; q < EPB * G2; q += S::NT) {
        const int el2 = q / G2, e = e0 + el2;
        if (e >= nelt) continue;
        const int w = q - el2 * G2;
        a.ghost[int64_t(e) * G2 + w] = double(blk[el2][ghost_ext<NH>(w / H2, w - (w / H2) * H2)]);
    }
}

template <typename T, int NH>
int launch_fdm_nh(const FdmArgs &a, const T *ops, const T *eps, hipStream_t st)
{
    using S = FdmShape<NH>;
    hipLaunchKernelGGL((k_fdm<T, NH>), dim3((a.nelt + S::EPB - 1) / S::EPB), dim3(S::NT), 0, st,
                       a, ops, eps);
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

template <typename T>
int launch_fdm(const FdmArgs &a, const T *ops, const T *eps, int nh, hipStream_t st)
{
    switch (nh) {
    case 3: return launch_fdm_nh<T, 3>(a, ops, eps, st);
    case 4: return launch_fdm_nh<T, 4>(a, ops, eps, st);
    case 5: return launch_fdm_nh<T, 5>(a, ops, eps, st);
    case 6: return launch_fdm_nh<T, 6>(a, ops, eps, st);
    case 7: return launch_fdm_nh<T, 7>(a, ops, eps, st);
    case 8: return launch_fdm_nh<T, 8>(a, ops, eps, st);
    case 9: return launch_fdm_nh<T, 9>(a, ops, eps, st);
    case 10: return launch_fdm_nh<T, 10>(a, ops, eps, st);
    case 11: return launch_fdm_nh<T, 11>(a, ops, eps, st);
    case 12: return launch_fdm_nh<T, 12>(a, ops, eps, st);
    case 13: return launch_fdm_nh<T, 13>(a, ops, eps, st);
    case 14: return launch_fdm_nh<T, 14>(a, ops, eps, st);
    case 15: return launch_fdm_nh<T, 15>(a, ops, eps, st);
    case 16: return launch_fdm_nh<T, 16>(a, ops, eps, st);
    default: return fail(NEKBONE_ERR_UNSUPPORTED, "hsmg: level size outside 3..16");
    }
}

// The rest of a Schwarz level after the local solves, in two passes: the
// overlap sum (src/hsmg.f:1081-1085), dssum on the level (:1089), the
// mask (:1090) and the square-rooted weights (h1mg_schwarz_wt3d2,
// :1144-1178):
//   v    = e_own + the neighbours' ghost results on the first interior
//          layer, x faces, then y, then z, each added as (own + theirs) - own
//          (the exchange of the (n+2)^3 ghost layer, :1071-1085)
//   t    = v at an unshared point; at a shared one v(member 0) + v(member 1)
//          + ... in the level numbering's member order, for every member
//   out  = (masked ? 0 : t) * sqrt(weight)
// k_swz_ovl computes v at every point, finishes the unshared points and
// leaves v in e_own at the shared ones; k_swz_gs sums the groups and
// finishes their members.  Every value rounds as in the sequence of the
// three steps.
struct SwzArgs {
    double *out, *eown;
    const double *ghost, *sw;
    const int32_t *gpart;
    const uint8_t *bnd, *shr;
    int shrbit;  // shr[m] & shrbit: point m is in a gs group
    int n;
    int b2, b4, b8, br;  // workgroups of each group class
    GsArgs G;
    const CgScalars *guard;  // may be null: return at once when guard->done
};

constexpr int kSwzPts = 4;  // points per lane in k_swz_ovl

// Ghost slots whose neighbour results point m receives, in the order they
// are added: direction x, y, z; within a direction the low side first
// (both sides only when NH = 3, where layer 1 is also layer NH-2); -1: none.
template <int NH, bool FINE>
__device__ __forceinline__ void swz_slots(int m, int (&slot)[NH == 3 ? 6 : 3])
{
    constexpr int N3 = NH * NH * NH, H2 = NH * NH, G2 = 6 * H2;
    const int e = m / N3, lm = m - e * N3;
    // the fine level is in the device point order, the others in the
    // reference order
    const int p = FINE ? surface_order_ref(NH, lm) : lm;
    const int i = p % NH, j = (p / NH) % NH, k = p / H2;
    const int c[3] = {i, j, k};
    const int t[3] = {j + NH * k, i + NH * k, i + NH * j};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if constexpr (NH == 3) {
            const bool on = c[d] == 1;
            slot[2 * d] = on ? e * G2 + (2 * d) * H2 + t[d] : -1;
            slot[2 * d + 1] = on ? e * G2 + (2 * d + 1) * H2 + t[d] : -1;
        } else {
            const int side = c[d] == 1 ? 0 : c[d] == NH - 2 ? 1 : -1;
            slot[d] = side < 0 ? -1 : e * G2 + (2 * d + side) * H2 + t[d];
        }
    }
}

// kSwzPts points per lane (lane + u * kRedThreads of the workgroup's
// range), every load of all of them issued before the neighbour gathers
// and those before the first store.
template <int NH, bool FINE>
__global__ void __launch_bounds__(kRedThreads) k_swz_ovl(SwzArgs a)
{
    if (a.guard && a.guard->done) return;
    constexpr int NC = NH == 3 ? 6 : 3;
    const int base = blockIdx.x * (kRedThreads * kSwzPts) + threadIdx.x;
    int slot[kSwzPts][NC], gp[kSwzPts][NC];
    double ev[kSwzPts], go[kSwzPts][NC], gv[kSwzPts][NC];
    uint8_t sh[kSwzPts];
#pragma unroll
    for (int u = 0; u < kSwzPts; ++u) {
        const int m = base + u * kRedThreads;
        const bool ok = m < a.n;
        swz_slots<NH, FINE>(ok ? m : 0, slot[u]);
        ev[u] = ok ? a.eown[m] : 0.0;
        sh[u] = ok ? a.shr[m] : 0;
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            if (!ok) slot[u][q] = -1;
            gp[u][q] = slot[u][q] >= 0 ? a.gpart[slot[u][q]] : -1;
            go[u][q] = slot[u][q] >= 0 ? a.ghost[slot[u][q]] : 0.0;
        }
    }
    // the mask and weight of the unshared points load with the gathers
    uint8_t bd[kSwzPts];
    double wt[kSwzPts];
#pragma unroll
    for (int u = 0; u < kSwzPts; ++u) {
        const int m = base + u * kRedThreads;
        const bool fin = m < a.n && !(sh[u] & a.shrbit);
        bd[u] = fin ? a.bnd[m] : 0;
        wt[u] = fin ? a.sw[m] : 0.0;
#pragma unroll
        for (int q = 0; q < NC; ++q) gv[u][q] = gp[u][q] >= 0 ? a.ghost[gp[u][q]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kSwzPts; ++u) {
        const int m = base + u * kRedThreads;
        if (m >= a.n) continue;
        double v = ev[u];
#pragma unroll
        for (int q = 0; q < NC; ++q)
            if (gp[u][q] >= 0) v = v + ((go[u][q] + gv[u][q]) - go[u][q]);
        if (sh[u] & a.shrbit) {
            a.eown[m] = v;
        } else {
            const double w = bd[u] ? 0.0 : v;
            a.out[m] = w * wt[u];
        }
    }
}

__device__ __forceinline__ void swz_store(const SwzArgs &a, int m, double t)
{
    const double v = a.bnd[m] ? 0.0 : t;
    a.out[m] = v * a.sw[m];
}

// Sum of one group and its members finished; every load before the
// first store.
template <int S>
__device__ __forceinline__ void swz_group(const SwzArgs &a, const int (&m)[S])
{
    double v[S], wt[S];
    uint8_t bd[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
        v[q] = a.eown[m[q]];
        bd[q] = a.bnd[m[q]];
        wt[q] = a.sw[m[q]];
    }
    double t = v[0];
#pragma unroll
    for (int q = 1; q < S; ++q) t += v[q];
#pragma unroll
    for (int q = 0; q < S; ++q) a.out[m[q]] = (bd[q] ? 0.0 : t) * wt[q];
}

// The level's gather-scatter classes (pairs, quads, octets, the rest),
// one group per lane, on the v that k_swz_ovl left at the shared points.
__global__ void __launch_bounds__(kRedThreads) k_swz_gs(SwzArgs a)
{
    if (a.guard && a.guard->done) return;
    const GsArgs &G = a.G;
    int blk = blockIdx.x;
    const int64_t g0 = int64_t(threadIdx.x);
    if (blk < a.b2) {
        const int64_t g = int64_t(blk) * kRedThreads + g0;
        if (g >= G.n2) return;
        int m[2];
        gs_members<2>(G.m2, G.m2r, G.r2, G.n2r, g, m);
        swz_group<2>(a, m);
    } else if ((blk -= a.b2) < a.b4) {
        const int64_t g = int64_t(blk) * kRedThreads + g0;
        if (g >= G.n4) return;
        int m[4];
        gs_members<4>(G.m4, G.m4r, G.r4, G.n4r, g, m);
        swz_group<4>(a, m);
    } else if ((blk -= a.b4) < a.b8) {
        const int64_t g = int64_t(blk) * kRedThreads + g0;
        if (g >= G.n8) return;
        int m[8];
        gs_members<8>(G.m8, nullptr, 1, 0, g, m);
        swz_group<8>(a, m);
    } else {
        blk -= a.b8;
        const int64_t g = int64_t(blk) * kRedThreads + g0;
        if (g >= G.nr) return;
        const int lo = G.roff[g], hi = G.roff[g + 1];
        double t = a.eown[G.ridx[lo]];
        for (int q = lo + 1; q < hi; ++q) t += a.eown[G.ridx[q]];
        for (int q = lo; q < hi; ++q) swz_store(a, G.ridx[q], t);
    }
}

template <int NH>
int launch_swz_nh(const SwzArgs &a, bool fine, hipStream_t st)
{
    const int nbo = int((int64_t(a.n) + kRedThreads * kSwzPts - 1) / (kRedThreads * kSwzPts));
    if (fine)
        hipLaunchKernelGGL((k_swz_ovl<NH, true>), dim3(nbo), dim3(kRedThreads), 0, st, a);
    else
        hipLaunchKernelGGL((k_swz_ovl<NH, false>), dim3(nbo), dim3(kRedThreads), 0, st, a);
    NB_HIP(hipGetLastError());
    const int nbg = a.b2 + a.b4 + a.b8 + a.br;
    if (nbg > 0) {
        hipLaunchKernelGGL(k_swz_gs, dim3(nbg), dim3(kRedThreads), 0, st, a);
        NB_HIP(hipGetLastError());
    }
    return NEKBONE_OK;
}

// Restriction of one element from the finer level F to the coarser C:
// rc = [J^T x J^T x J^T] (c_F .* rf) (h1mg_do_wt then h1mg_tnsr1_3d);
// with MASK the coarse result is written masked; with MIN the fine input
// is read masked by bndf (the finest level's mask, which that level's
// Schwarz step applies to rf in place while this kernel runs beside it:
// either value read gives the masked one).
template <bool MASK, bool MIN = false>
__global__ void __launch_bounds__(kT)
    k_rstr(double *__restrict__ rc, const double *__restrict__ rf, const double *__restrict__ cf,
           const int32_t *__restrict__ posf, const double *__restrict__ jh,
           const double *__restrict__ jht, const uint8_t *__restrict__ bndc, int nc, int nf,
           const CgScalars *__restrict__ guard, const uint8_t *__restrict__ bndf)
{
    if (guard && guard->done) return;
    extern __shared__ double sm[];
    const int f3 = nf * nf * nf, c3 = nc * nc * nc;
    double *U = sm, *W1 = sm + f3, *W2 = W1 + nc * nf * nf, *J = W2 + nc * nc * nf, *Jt = J + nc * nf;
    const int64_t e = blockIdx.x;
    for (int p = threadIdx.x; p < f3; p += blockDim.x) {
        const int64_t m = e * f3 + posf[p];
        const double v = MIN && bndf[m] ? 0.0 : rf[m];
        U[p] = v * cf[m];
    }
    for (int p = threadIdx.x; p < nc * nf; p += blockDim.x) {
        J[p] = jh[p];
        Jt[p] = jht[p];
    }
    __syncthreads();
    tx(W1, U, Jt, nc, nf, nf * nf);          // A = jht (nc x nf)
    __syncthreads();
    ty(W2, W1, J, nc, nc, nf, nf);            // Bt = jh (nf x nc)
    __syncthreads();
    tz(U, W2, J, nc * nc, nc, nf);            // Ct = jh
    __syncthreads();
    for (int p = threadIdx.x; p < c3; p += blockDim.x) {
        const int64_t m = e * c3 + p;
        rc[m] = (MASK && bndc[m]) ? 0.0 : U[p];
    }
}

// Interpolation-add of one element: uf += [J x J x J] uc
// (h1mg_intp then the add of h1mg_solve); with MASK the coarse input is
// masked first (the coarse solve's final mask, :1011).
// RTZ (the finest level's interpolation, which completes z inside a CG):
// also the element's part of rtz = glsc3(r, c, z) (src/cg.f:52) into
// part[e], so no separate pass reads r, c and z again.
template <bool MASK, bool RTZ = false>
__global__ void __launch_bounds__(kT)
    k_intp(double *__restrict__ uf, const int32_t *__restrict__ posf,
           const double *__restrict__ uc, const uint8_t *__restrict__ bndc,
           const double *__restrict__ jh, const double *__restrict__ jht, int nc, int nf,
           const CgScalars *__restrict__ guard, const double *__restrict__ rr = nullptr,
           const double *__restrict__ cc = nullptr, double *__restrict__ part = nullptr)
{
    if (guard && guard->done) return;
    extern __shared__ double sm[];
    const int f3 = nf * nf * nf, c3 = nc * nc * nc;
    double *U = sm, *W1 = sm + f3, *W2 = W1 + nf * nc * nc, *J = W2 + nf * nf * nc, *Jt = J + nc * nf;
    const int64_t e = blockIdx.x;
    for (int p = threadIdx.x; p < c3; p += blockDim.x) {
        const int64_t m = e * c3 + p;
        U[p] = (MASK && bndc[m]) ? 0.0 : uc[m];
    }
    for (int p = threadIdx.x; p < nc * nf; p += blockDim.x) {
        J[p] = jh[p];
        Jt[p] = jht[p];
    }
    __syncthreads();
    tx(W1, U, J, nf, nc, nc * nc);            // A = jh (nf x nc)
    __syncthreads();
    ty(W2, W1, Jt, nf, nf, nc, nc);           // Bt = jht (nc x nf)
    __syncthreads();
    tz(U, W2, Jt, nf * nf, nf, nc);           // Ct = jht
    __syncthreads();
    double acc = 0.0;
    for (int p = threadIdx.x; p < f3; p += blockDim.x) {
        const int64_t m = e * f3 + posf[p];
        const double v = uf[m] + U[p];
        uf[m] = v;
        if constexpr (RTZ) acc = acc + rr[m] * cc[m] * v;
    }
    if constexpr (RTZ) {
        double dummy = 0.0;
        block_sum2<kTE>(acc, dummy);
        if (threadIdx.x == 0) part[e] = acc;
    }
}

int grid(int64_t n) { return int((n + kT - 1) / kT); }

// One Schwarz level: out = W^1/2 mask(QQ^T(ovl(fdm(mask(r))))); r masked in place.
int schwarz(nekbone_problem *P, HsmgLevel &L, double *out, double *r, const CgScalars *guard,
            hipStream_t st)
{
    const FdmArgs a{r,       L.d_bnd,  L.d_pos,   L.d_gsrc, L.d_gin,
                    L.d_eop, L.d_eown, L.d_ghost, P->nelt,  guard};
    if (int rc = P->hsmg_fp32 ? launch_fdm<float>(a, L.d_ops32, L.d_eps32, L.nh, st)
                              : launch_fdm<double>(a, L.d_ops, L.d_eps, L.nh, st))
        return rc;
    if (int64_t(6) * L.nh * L.nh * P->nelt > INT32_MAX || L.n > INT32_MAX)
        return fail(NEKBONE_ERR_UNSUPPORTED, "hsmg: level above 2^31 points or ghost slots");
    SwzArgs s{};
    s.out = out;
    s.eown = L.d_eown;
    s.ghost = L.d_ghost;
    s.sw = L.d_sw;
    s.gpart = L.d_gpart;
    s.bnd = L.d_bnd;
    s.n = int(L.n);
    s.guard = guard;
    if (L.fine) {
        // the fine level is the problem's own numbering (device order)
        s.G = gs_args(P);
        s.shr = P->d_flag;
        s.shrbit = kFlagShared;
    } else {
        s.G = gs_args_map(L.gs);
        s.shr = L.d_shr;
        s.shrbit = 1;
    }
    auto nb = [](int64_t k) { return int((k + kRedThreads - 1) / kRedThreads); };
    s.b2 = nb(s.G.n2);
    s.b4 = nb(s.G.n4);
    s.b8 = nb(s.G.n8);
    s.br = nb(s.G.nr);
    switch (L.nh) {
    case 3: return launch_swz_nh<3>(s, L.fine, st);
    case 4: return launch_swz_nh<4>(s, L.fine, st);
    case 5: return launch_swz_nh<5>(s, L.fine, st);
    case 6: return launch_swz_nh<6>(s, L.fine, st);
    case 7: return launch_swz_nh<7>(s, L.fine, st);
    case 8: return launch_swz_nh<8>(s, L.fine, st);
    case 9: return launch_swz_nh<9>(s, L.fine, st);
    case 10: return launch_swz_nh<10>(s, L.fine, st);
    case 11: return launch_swz_nh<11>(s, L.fine, st);
    case 12: return launch_swz_nh<12>(s, L.fine, st);
    case 13: return launch_swz_nh<13>(s, L.fine, st);
    case 14: return launch_swz_nh<14>(s, L.fine, st);
    case 15: return launch_swz_nh<15>(s, L.fine, st);
    case 16: return launch_swz_nh<16>(s, L.fine, st);
    default: return fail(NEKBONE_ERR_UNSUPPORTED, "hsmg: level size outside 3..16");
    }
}

int rstr(nekbone_problem *P, HsmgLevel &C, HsmgLevel &F, double *rc, const double *rf, bool mask,
         const CgScalars *guard, bool maskin = false)
{
    const int nc = C.nh, nf = F.nh;
    const size_t lds =
        sizeof(double) * size_t(nf * nf * nf + nc * nf * nf + nc * nc * nf + 2 * nc * nf);
#define NB_RSTR(MASK, MIN)                                                                  \
    hipLaunchKernelGGL((k_rstr<MASK, MIN>), dim3(P->nelt), dim3(kTE), lds, stream(), rc, rf, \
                       static_cast<const double *>(F.d_c),                                   \
                       static_cast<const int32_t *>(F.d_pos),                                \
                       static_cast<const double *>(C.d_jh),                                  \
                       static_cast<const double *>(C.d_jht),                                 \
                       static_cast<const uint8_t *>(C.d_bnd), nc, nf, guard,                 \
                       static_cast<const uint8_t *>(F.d_bnd))
    if (mask && maskin)
        NB_RSTR(true, true);
    else if (mask)
        NB_RSTR(true, false);
    else if (maskin)
        NB_RSTR(false, true);
    else
        NB_RSTR(false, false);
#undef NB_RSTR
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

// rtz_part (may be null): k_intp's RTZ form, the parts of glsc3(r, c, uf)
int intp(nekbone_problem *P, HsmgLevel &C, HsmgLevel &F, double *uf, const double *uc, bool mask,
         const CgScalars *guard, double *rtz_part = nullptr, const double *r = nullptr)
{
    const int nc = C.nh, nf = F.nh;
    const size_t lds =
        sizeof(double) * size_t(nf * nf * nf + nf * nc * nc + nf * nf * nc + 2 * nc * nf);
#define NB_INTP(MASK, RTZ)                                                                  \
    hipLaunchKernelGGL((k_intp<MASK, RTZ>), dim3(P->nelt), dim3(kTE), lds, stream(), uf,      \
                       static_cast<const int32_t *>(F.d_pos), uc,                            \
                       static_cast<const uint8_t *>(C.d_bnd),                                \
                       static_cast<const double *>(C.d_jh),                                  \
                       static_cast<const double *>(C.d_jht), nc, nf, guard, r,               \
                       static_cast<const double *>(P->d_c), rtz_part)
    if (mask && rtz_part)
        NB_INTP(true, true);
    else if (mask)
        NB_INTP(true, false);
    else if (rtz_part)
        NB_INTP(false, true);
    else
        NB_INTP(false, false);
#undef NB_INTP
    NB_HIP(hipGetLastError());
    return NEKBONE_OK;
}

}  // namespace

// h1mg_solve (src/hsmg.f:960-1034): z = M^-1 r on the finest level
// (device point order); r is masked in place.
int launch_hsmg_apply(nekbone_problem *P, double *z, double *r, const CgScalars *guard,
                      double *rtz_part)
{
    Hsmg *M = P->mg;
    if (!M) return fail(NEKBONE_ERR_STATE, "hsmg: not set up");
    const int lmax = M->lmax;
    HsmgLevel *L = M->L;
    if (int rc = timed_begin(P, 2)) return rc;
    // Additive Schwarz: each level's Schwarz step depends only on that
    // level's r, and the coarser levels only on the restricted r.  The
    // finest level's step runs on side stream 0, forked after the first
    // restriction (so that the small kernels of the chain are not queued
    // behind its local solves), the last middle level's on side stream 1,
    // forked after that level's r is summed; each is joined before the
    // interpolation that adds into its output.  A Schwarz step masks its
    // r in place as it reads it (the reference's side effect), so the
    // restriction that reads that r beside it masks as it reads (maskin):
    // every value is the serial order's.
    const double *rcur = r;  // r := rhs, masked by the Schwarz step
    bool fine_forked = false, mid_forked = false;
    auto fork_fine = [&]() -> int {
        if (fine_forked) return NEKBONE_OK;
        fine_forked = true;
        if (int rc = fork_side(0)) return rc;
        return schwarz(P, L[lmax], z, r, guard, side_stream(0));
    };
    for (int l = lmax - 1; l >= 2; --l) {
        if (int rc = rstr(P, L[l], L[l + 1], L[l].d_r, rcur, false, guard, l == lmax - 1))
            return rc;
        if (int rc = fork_fine()) return rc;
        if (int rc = launch_gs_map(L[l].gs, L[l].d_r, guard)) return rc;
        if (l == 2) {
            if (int rc = fork_side(1)) return rc;
            if (int rc = schwarz(P, L[l], L[l].d_e, L[l].d_r, guard, side_stream(1))) return rc;
            mid_forked = true;
        } else if (int rc = schwarz(P, L[l], L[l].d_e, L[l].d_r, guard, stream())) {
            return rc;
        }
        rcur = L[l].d_r;
    }
    // coarse level: e1 = mask(QQ^T mask(J^T r2)) (hsmg_coarse_solve, :945-957),
    // or with the direct coarse solve e1 = Q A_c^-1 Q^T mask(J^T r2)
    if (int rc = rstr(P, L[1], L[2], L[1].d_e, rcur, true, guard, true)) return rc;
    if (int rc = fork_fine()) return rc;  // two levels: the coarse restriction was the first
    if (int rc = launch_gs_map(L[1].gs, L[1].d_e, guard)) return rc;
    if (P->hsmg_coarse)
        if (int rc = hsmg_coarse_apply(P, L[1].d_e, guard)) return rc;
    for (int l = 2; l <= lmax - 1; ++l) {
        if (l == 2 && mid_forked)
            if (int rc = join_side(1)) return rc;
        if (int rc = intp(P, L[l - 1], L[l], L[l].d_e, L[l - 1].d_e, l - 1 == 1, guard)) return rc;
    }
    if (int rc = join_side(0)) return rc;
    if (int rc = intp(P, L[lmax - 1], L[lmax], z, L[lmax - 1].d_e, lmax - 1 == 1, guard,
                      rtz_part, r))
        return rc;
    return timed_end(P, 2);
}

}  // namespace nb

using namespace nb;

extern "C" int nekbone_setup_hsmg(nekbone_problem *P)
{
    if (int rc = check(P)) return rc;
    if (int rc = hsmg_build(P)) {
        hsmg_free(P);
        return rc;
    }
    NB_HIP(hipStreamSynchronize(stream()));
    return NEKBONE_OK;
}

extern "C" int nekbone_set_hsmg_fp32(nekbone_problem *P, int on)
{
    if (int rc = check(P)) return rc;
    P->hsmg_fp32 = on ? 1 : 0;
    return NEKBONE_OK;
}

extern "C" int nekbone_hsmg_solve(nekbone_problem *P, double *z, double *r)
{
    if (int rc = check(P)) return rc;
    if (!P->mg) return fail(NEKBONE_ERR_STATE, "hsmg_solve: nekbone_setup_hsmg not called");
    if (!z || !r) return fail(NEKBONE_ERR_ARG, "hsmg_solve: null array");
    if (int rc = upload_dev(P, P->d_r, r)) return rc;
    if (int rc = launch_hsmg_apply(P, P->mg->d_z, P->d_r)) return rc;
    if (int rc = download_ref(P, z, P->mg->d_z)) return rc;
    return download_ref(P, r, P->d_r);
}
