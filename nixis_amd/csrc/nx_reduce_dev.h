// nx_reduce_dev.h -- device helpers for order-preserving min/max reductions.
#pragma once

#include <hip/hip_runtime.h>

namespace nx {

// Monotone map double -> uint64 (negative values reversed), so unsigned
// integer min/max atomics order doubles correctly.
__device__ __forceinline__ unsigned long long dkey(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double dunkey(unsigned long long k)
{
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

// Block-wide (min, max) of keys; the result is valid in thread 0.
__device__ __forceinline__ void block_minmax(unsigned long long &lo, unsigned long long &hi)
{
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long a = __shfl_xor(lo, off, 64), b = __shfl_xor(hi, off, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    __shared__ unsigned long long rl[16], rh[16];
    if ((threadIdx.x & 63) == 0) {
        rl[threadIdx.x >> 6] = lo;
        rh[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            lo = rl[w] < lo ? rl[w] : lo;
            hi = rh[w] > hi ? rh[w] : hi;
        }
    }
}

}  // namespace nx
