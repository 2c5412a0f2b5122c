// Device-side helpers shared by the kernel files.
#pragma once

#include "nb_internal.h"

namespace nb {

constexpr int kRedThreads = 256;
constexpr int kMaxRedBlocks = 1024;

// Fixed-shape block tree reduction of two values over NT threads
// (deterministic: the combination order depends only on thread indices).
template <int NT = kRedThreads>
__device__ __forceinline__ void block_sum2(double &v0, double &v1)
{
    __shared__ double s0[NT], s1[NT];
    const int t = threadIdx.x;
    s0[t] = v0;
    s1[t] = v1;
    __syncthreads();
#pragma unroll
    for (int h = NT / 2; h > 0; h >>= 1) {
        if (t < h) {
            s0[t] = s0[t] + s0[t + h];
            s1[t] = s1[t] + s1[t + h];
        }
        __syncthreads();
    }
    v0 = s0[0];
    v1 = s1[0];
}

}  // namespace nb
