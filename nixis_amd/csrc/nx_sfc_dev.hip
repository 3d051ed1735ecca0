// nx_sfc_dev.hip -- space-filling-curve order and neighbour layout on the device.
//
// The same results as the host versions in nx_mesh.cpp, bit for bit:
//   * sfc_order_dev: bounding box (order-preserving min/max keys), 3D Morton
//     key of the coordinates quantised to 21 bits per axis with the host's
//     FP64 formula, then a stable radix sort of (key, index) pairs, so ties
//     keep index order exactly as the host's stable LSD sort;
//   * ell_from_adjacency_dev: slot-major neighbour table col[6][V] and FP64
//     edge lengths elen[6][V] (erosion.py:38-40) in the new numbering, a
//     pad slot pointing at the vertex itself with length 0.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "nx_ctx.h"
#include "nx_reduce_dev.h"

namespace nx {

namespace {

inline size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

__device__ __forceinline__ uint64_t spread21(uint64_t x)
{
    x &= 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}

struct SfcLayout {
    size_t box, key_in, key_out, idx_in, temp, total;
};

SfcLayout sfc_layout(int64_t V, size_t temp_bytes)
{
    SfcLayout L;
    size_t o = 0;
    L.box = o;
    o += 256;
    L.key_in = o;
    o += a256(8 * (size_t)V);
    L.key_out = o;
    o += a256(8 * (size_t)V);
    L.idx_in = o;
    o += a256(8 * (size_t)V);
    L.temp = o;
    o += a256(temp_bytes);
    L.total = o;
    return L;
}

size_t sort_temp_bytes(int64_t V)
{
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint64_t *)nullptr,
                                             (uint64_t *)nullptr, (const int64_t *)nullptr,
                                             (int64_t *)nullptr, (int)V, 0, 63);
    return bytes;
}

}  // namespace

__global__ void k_box_init(unsigned long long *box)
{
    if (threadIdx.x < 3) {
        box[2 * threadIdx.x] = ~0ull;
        box[2 * threadIdx.x + 1] = 0ull;
    }
}

__global__ __launch_bounds__(256) void k_box(const double *__restrict__ verts, int64_t V,
                                             unsigned long long *__restrict__ box)
{
    for (int k = 0; k < 3; ++k) {
        unsigned long long lo = ~0ull, hi = 0ull;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V;
             i += (int64_t)gridDim.x * blockDim.x) {
            const unsigned long long q = dkey(verts[3 * i + k]);
            lo = q < lo ? q : lo;
            hi = q > hi ? q : hi;
        }
        block_minmax(lo, hi);
        if (threadIdx.x == 0) {
            __hip_atomic_fetch_min(&box[2 * k], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_max(&box[2 * k + 1], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_morton(const double *__restrict__ verts, int64_t V,
                                                const unsigned long long *__restrict__ box,
                                                uint64_t *__restrict__ key,
                                                int64_t *__restrict__ idx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    uint64_t q[3];
    for (int k = 0; k < 3; ++k) {
        const double lo = dunkey(box[2 * k]), hi = dunkey(box[2 * k + 1]);
        const double span = hi - lo;
        const double t = span > 0 ? (verts[3 * i + k] - lo) / span : 0.0;
        q[k] = (uint64_t)fmin(2097151.0, fmax(0.0, t * 2097151.0));
    }
    key[i] = spread21(q[0]) | spread21(q[1]) << 1 | spread21(q[2]) << 2;
    idx[i] = i;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// inv[order[i]] = i, rejecting out-of-range and repeated entries.
__global__ __launch_bounds__(256) void k_inverse(const int64_t *__restrict__ order, int64_t V,
                                                 int32_t *__restrict__ inv, int *__restrict__ err)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const int64_t o = order[i];
    if (o < 0 || o >= V) {
        __hip_atomic_fetch_or(err, NX_ERR_MESH_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    if (atomicExch(&inv[o], (int32_t)i) != -1)
        __hip_atomic_fetch_or(err, NX_ERR_MESH_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(256) void k_ell(const int32_t *__restrict__ adj, int64_t V,
                                             const double *__restrict__ verts,
                                             const int64_t *__restrict__ order,
                                             const int32_t *__restrict__ inv,
                                             int32_t *__restrict__ col, double *__restrict__ elen,
                                             int *__restrict__ err)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const int64_t old = order ? order[i] : i;
    if (old < 0 || old >= V) {
        __hip_atomic_fetch_or(err, NX_ERR_MESH_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const double *p = verts + 3 * old;
    bool bad = false;
    for (int k = 0; k < 6; ++k) {
        const int32_t n = adj[old * 6 + k];
        const int64_t e = (int64_t)k * V + i;
        if (n < -1 || n >= V) {
            bad = true;
            col[e] = (int32_t)i;
            elen[e] = 0.0;
        } else if (n == -1) {
            col[e] = (int32_t)i;
            elen[e] = 0.0;
        } else {
            const double *q = verts + 3 * (int64_t)n;
            const double a = p[0] - q[0], b = p[1] - q[1], c = p[2] - q[2];
            col[e] = order ? inv[n] : n;
            elen[e] = sqrt(a * a + b * b + c * c);
        }
    }
    if (bad) __hip_atomic_fetch_or(err, NX_ERR_MESH_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

static unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

size_t sfc_workspace_bytes(int64_t V)
{
    if (V <= 0 || V > INT32_MAX) return 256;
    return sfc_layout(V, sort_temp_bytes(V)).total;
}

int sfc_order_dev(nx_ctx *ctx, const double *verts, int64_t V, int64_t *order, void *ws,
                  hipStream_t stream)
{
    if (!ctx || !verts || !order || !ws || V <= 0 || V > INT32_MAX)
        return set_error(NX_EINVAL, "sfc order needs verts, order, a workspace and 0 < V < 2^31");
    if (((uintptr_t)ws & 255) != 0) return set_error(NX_EINVAL, "workspace must be 256-byte aligned");
    NX_HIP(hipSetDevice(ctx->device));
    size_t temp = sort_temp_bytes(V);
    const SfcLayout L = sfc_layout(V, temp);
    char *w = (char *)ws;
    auto *box = (unsigned long long *)(w + L.box);
    auto *kin = (uint64_t *)(w + L.key_in);
    auto *kout = (uint64_t *)(w + L.key_out);
    auto *iin = (int64_t *)(w + L.idx_in);
    hipLaunchKernelGGL(k_box_init, dim3(1), dim3(64), 0, stream, box);
    hipLaunchKernelGGL(k_box, dim3(std::min<unsigned>(blocks_for(V), 2048u)), dim3(256), 0, stream,
                       verts, V, box);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(V)), dim3(256), 0, stream, verts, V, box, kin, iin);
    NX_HIP(hipGetLastError());
    NX_HIP(hipcub::DeviceRadixSort::SortPairs(w + L.temp, temp, kin, kout, iin, order, (int)V, 0,
                                              63, stream));
    return NX_OK;
}

size_t ell_workspace_bytes(int64_t V) { return V > 0 ? a256(4 * (size_t)V) : 256; }

int ell_from_adjacency_dev(nx_ctx *ctx, const int32_t *adj, int64_t V, const double *verts,
                           const int64_t *order, int32_t *col, double *elen, void *ws,
                           hipStream_t stream)
{
    if (!ctx || !adj || !verts || !col || !elen || V <= 0 || V > INT32_MAX || (order && !ws))
        return set_error(NX_EINVAL, "neighbour layout needs adj, verts, col, elen (and a workspace with an order)");
    NX_HIP(hipSetDevice(ctx->device));
    int32_t *inv = (int32_t *)ws;
    if (order) {
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(V)), dim3(256), 0, stream, inv, V, -1);
        hipLaunchKernelGGL(k_inverse, dim3(blocks_for(V)), dim3(256), 0, stream, order, V, inv,
                           ctx->d_err);
    }
    hipLaunchKernelGGL(k_ell, dim3(blocks_for(V)), dim3(256), 0, stream, adj, V, verts, order,
                       order ? inv : nullptr, col, elen, ctx->d_err);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

}  // namespace nx
