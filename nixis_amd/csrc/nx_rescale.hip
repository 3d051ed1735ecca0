// nx_rescale.hip -- util.py:110-140 (rescale, standard mode) on the device.
//
// Two passes: a min/max reduction (per-block reduction, then one atomic per
// block on an order-preserving integer encoding of the doubles), then the
// map out = ((x - min) / (max - min)) * (upper - lower) + lower with the
// reference's FP64 operation order (IEEE division, no contraction), so the
// result is bit-identical to the reference's for the same input.
#include <hip/hip_runtime.h>

#include <atomic>

#include "nx_ctx.h"
#include "nx_reduce_dev.h"

namespace nx {

__global__ void k_minmax_init(unsigned long long *keys)
{
    if (threadIdx.x == 0) {
        keys[0] = ~0ull;  // running min
        keys[1] = 0ull;   // running max
    }
}

__global__ __launch_bounds__(256) void k_minmax(const double *__restrict__ x, int64_t n,
                                                unsigned long long *__restrict__ keys)
{
    unsigned long long lo = ~0ull, hi = 0ull;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long k = dkey(x[i]);
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long a = __shfl_xor(lo, off, 64), b = __shfl_xor(hi, off, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    __shared__ unsigned long long rl[4], rh[4];
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
        __hip_atomic_fetch_min(&keys[0], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_max(&keys[1], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// x and out may alias (in-place rescale): each element is read and then
// written by the same thread after the reduction has finished.
__global__ __launch_bounds__(256) void k_rescale_map(const double *x, int64_t n,
                                                     const unsigned long long *__restrict__ keys,
                                                     double lower, double upper,
                                                     double *out,
                                                     double *__restrict__ minmax,
                                                     int *__restrict__ err)
{
    const double x_min = dunkey(keys[0]), x_max = dunkey(keys[1]);
    const double x_range = x_max - x_min;
    const double new_range = upper - lower;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (minmax) {
            minmax[0] = x_min;
            minmax[1] = x_max;
        }
        if (x_range == 0.0)
            __hip_atomic_fetch_or(err, NX_ERR_DIV0_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = ((x[i] - x_min) / x_range) * new_range + lower;  // util.py:139
}

int rescale(nx_ctx *ctx, const double *x, int64_t n, double lower, double upper, double *out,
            double *minmax, hipStream_t stream)
{
    if (!ctx) return set_error(NX_EINVAL, "null context");
    if (n <= 0 || !x || !out) return set_error(NX_EINVAL, "rescale needs n > 0 and device arrays");
    NX_HIP(hipSetDevice(ctx->device));
    // Reduction keys: one 16-byte slot per call from a ring in the context;
    // 256 concurrent rescale calls on one context would be needed to reuse a
    // slot still in flight.
    unsigned long long *keys = ctx->d_keys + 2 * (ctx->key_slot.fetch_add(1) % kKeySlots);
    const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cus * 4);
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, stream, keys);
    hipLaunchKernelGGL(k_minmax, dim3(grid), dim3(256), 0, stream, x, n, keys);
    hipLaunchKernelGGL(k_rescale_map, dim3(grid), dim3(256), 0, stream, x, n, keys, lower, upper,
                       out, minmax, ctx->d_err);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

}  // namespace nx
