// Device-side helpers shared by the kernel files.
#pragma once

#include "nb_internal.h"

namespace nb {

constexpr int kRedThreads = 256;
constexpr int kMaxRedBlocks = 1024;

// Logical block of hardware workgroup i of nblk, walked upward (rev = 0)
// or downward (rev = 1).  Workgroups start roughly in index order, so
// consecutive CG kernels that alternate the direction each start on the
// data their predecessor touched last, which the memory-side Infinity
// Cache may still hold.  Per-block results are stored under the logical
// block, so the direction changes no value.
__device__ __forceinline__ int sweep_block(int i, int nblk, int rev)
{
    return rev ? nblk - 1 - i : i;
}

// Sum over the 64 lanes of a wave by a butterfly of cross-lane moves:
// every lane ends with the same value, combined in an order fixed by the
// lane indices.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
    return v;
}

// Fixed-shape block reduction of two values over NT threads (a multiple
// of 64): wave butterflies, then the wave sums in wave order.  The result
// depends only on thread indices (deterministic); every thread returns
// it.  One barrier for the exchange, one so that the helper can be used
// again in the same kernel.
template <int NT = kRedThreads>
__device__ __forceinline__ void block_sum2(double &v0, double &v1)
{
    static_assert(NT % 64 == 0, "block of whole waves");
    constexpr int NW = NT / 64;
    __shared__ double s0[NW], s1[NW];
    const int t = threadIdx.x;
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    if ((t & 63) == 0) {
        s0[t >> 6] = v0;
        s1[t >> 6] = v1;
    }
    __syncthreads();
    double a = s0[0], b = s1[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) {
        a = a + s0[w];
        b = b + s1[w];
    }
    __syncthreads();
    v0 = a;
    v1 = b;
}

// Members of group g of a class of size S: the first nrun groups are in
// run form (runs of R; group t of a run has the run's first members + t),
// the rest in the member list.
template <int S>
__device__ __forceinline__ void gs_members(const int32_t *__restrict__ list,
                                           const int32_t *__restrict__ runs, int R, int64_t nrun,
                                           int64_t g, int (&m)[S])
{
    const int32_t *mem;
    int t = 0;
    if (g < nrun) {
        const int64_t run = g / R;
        t = int(g - run * R);
        mem = runs + S * run;
    } else {
        mem = list + S * (g - nrun);
    }
    if constexpr (S == 2) {
        const int2 v = *reinterpret_cast<const int2 *>(mem);
        m[0] = v.x + t, m[1] = v.y + t;
    } else {
#pragma unroll
        for (int h = 0; h < S / 4; ++h) {
            const int4 v = reinterpret_cast<const int4 *>(mem)[h];
            m[4 * h] = v.x + t, m[4 * h + 1] = v.y + t, m[4 * h + 2] = v.z + t,
                  m[4 * h + 3] = v.w + t;
        }
    }
}

}  // namespace nb
