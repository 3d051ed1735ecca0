// nx_erosion.hip -- erosion Jacobi steps over the icosphere neighbour table.
//
// erosion.py:171-192 (erode_terrain3) adds rain to every vertex and runs
// erosion.py:197-279 (erosion_iteration3) once per iteration; the latter
// copies h/w/s, gathers every vertex's neighbours, and copies the results
// back.  Here the fields stay on the device in a double-buffered workspace
// and one kernel launch is one Jacobi step:
//   * fields are float-float (hi, lo) pairs: height+water packed as float4
//     (one 16-byte gather per neighbour), sediment as float2;
//   * neighbours come from the slot-major table col[6][V] with float-float
//     edge lengths elen[6][V] (coalesced per slot);
//   * the water/sediment magnitudes grow about 2x per step (the vertex's own
//     water is re-added each step, erosion.py:215,264), so all fields share a
//     per-planet power-of-two exponent E.  A step whose outputs exceed 2^64
//     anywhere raises a flag (one wave ballot, then a plain store -- no
//     read-modify-write atomics on the hot path); the next step then scales
//     its inputs by 2^-64 (exact: the update is homogeneous in (h, w, s)
//     and branch decisions only compare heights) and adds the rain at scale
//     2^-E.  Magnitudes never shrink across the planet, so no scale-up path
//     is needed; loading normalises the largest input to [1, 2);
//   * the next step's rain is folded into this step's water output.
// Workspace: [meta 256 B][hw 2 x V x float4][s 2 x V x float2].
#include <hip/hip_runtime.h>

#include "nx_ctx.h"
#include "nx_erosion_dev.h"

namespace nx {

struct ErosionBufs {
    ErosionMeta *meta;
    float4 *hw[2];
    float2 *s[2];
};

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t erosion_state_bytes(int64_t V)
{
    if (V <= 0) return 256;
    return 256 + 2 * align256((size_t)V * sizeof(float4)) + 2 * align256((size_t)V * sizeof(float2));
}

static ErosionBufs bufs(void *state, int64_t V)
{
    char *p = (char *)state;
    ErosionBufs b;
    b.meta = (ErosionMeta *)p;
    p += 256;
    for (int k = 0; k < 2; ++k) {
        b.hw[k] = (float4 *)p;
        p += align256((size_t)V * sizeof(float4));
    }
    for (int k = 0; k < 2; ++k) {
        b.s[k] = (float2 *)p;
        p += align256((size_t)V * sizeof(float2));
    }
    return b;
}

__global__ __launch_bounds__(256) void k_erode_step(ErosionMeta *__restrict__ meta,
                                                    float4 *__restrict__ hw0,
                                                    float4 *__restrict__ hw1,
                                                    float2 *__restrict__ s0,
                                                    float2 *__restrict__ s1, int64_t V,
                                                    const int32_t *__restrict__ col,
                                                    const float2 *__restrict__ elen,
                                                    int64_t nblocks, int64_t k, double rain_next)
{
    // Step index from the workspace, so replaying captured launches works.
    const int64_t it = meta->it + k;
    const float4 *hw_src = (it & 1) ? hw1 : hw0;
    const float2 *s_src = (it & 1) ? s1 : s0;
    float4 *hw_dst = (it & 1) ? hw0 : hw1;
    float2 *s_dst = (it & 1) ? s0 : s1;
    const int cur = (int)(it % 3), prev = (int)((it + 2) % 3), nxt = (int)((it + 1) % 3);
    const int sh = meta->big[prev] ? 64 : 0;
    const int E = meta->E[prev] + sh;
    const float sc = ldexpf(1.0f, -sh);
    const int64_t tile = xcd_tile(blockIdx.x, nblocks);
    if (tile == 0 && threadIdx.x == 0) {
        meta->E[cur] = E;
        meta->big[nxt] = 0u;
    }
    float m = 0.0f;
    const int64_t i = tile * blockDim.x + threadIdx.x;
    if (tile < nblocks && i < V) {
        const ff rain = ff_ldexp(ff_from_double(rain_next), -E);
        m = erode_vertex_step(i, V, hw_src, s_src, hw_dst, s_dst, col, elen, sc, rain);
    }
    flag_big(m > kBig, &meta->big[cur]);
}

// Add rain (scaled by the current exponent) to every vertex's water in the
// current buffer: the reference's water += rain before the first step.
__global__ __launch_bounds__(256) void k_add_rain(ErosionMeta *__restrict__ meta,
                                                  float4 *__restrict__ hw0,
                                                  float4 *__restrict__ hw1, int64_t V,
                                                  double rain)
{
    const int64_t it = meta->it;
    float4 *hw = (it & 1) ? hw1 : hw0;
    const int prev = (int)((it + 2) % 3);
    const ff r = ff_ldexp(ff_from_double(rain), -meta->E[prev]);
    float m = 0.0f;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) {
        float4 v = hw[i];
        const ff w = ff_add(ff{v.z, v.w}, r);
        v.z = w.hi;
        v.w = w.lo;
        hw[i] = v;
        m = fabsf(w.hi);
    }
    flag_big(m > kBig, &meta->big[prev]);
}

// Advance the step counter after `iters` launches (one thread).
__global__ void k_advance(ErosionMeta *meta, int64_t iters) { meta->it += iters; }

// ---- load / store --------------------------------------------------------

__global__ __launch_bounds__(256) void k_load_max(ErosionMeta *__restrict__ meta,
                                                  const double *__restrict__ h,
                                                  const double *__restrict__ w,
                                                  const double *__restrict__ s, int64_t V)
{
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V;
         i += (int64_t)gridDim.x * blockDim.x) {
        m = fmax(m, fabs(h[i]));
        if (w) m = fmax(m, fabs(w[i]));
        if (s) m = fmax(m, fabs(s[i]));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = fmax(m, red[k]);
        __hip_atomic_fetch_max(&meta->load_max, (unsigned long long)__double_as_longlong(m),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ __launch_bounds__(256) void k_load(ErosionMeta *__restrict__ meta,
                                              float4 *__restrict__ hw, float2 *__restrict__ sd,
                                              const double *__restrict__ h,
                                              const double *__restrict__ w,
                                              const double *__restrict__ s, int64_t V)
{
    const double mx = __longlong_as_double((long long)meta->load_max);
    const int E = (mx > 0.0 && isfinite(mx)) ? ilogb(mx) : 0;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        meta->E[2] = E;  // exponent "after step -1"; max input now in [1, 2)
        meta->big[2] = 0u;
        meta->big[0] = 0u;
        meta->it = 0;
    }
    if (i < V) {
        const ff hh = ff_from_double(ldexp(h[i], -E));
        const ff ww = w ? ff_from_double(ldexp(w[i], -E)) : ff{0.0f, 0.0f};
        const ff ss = s ? ff_from_double(ldexp(s[i], -E)) : ff{0.0f, 0.0f};
        hw[i] = make_float4(hh.hi, hh.lo, ww.hi, ww.lo);
        sd[i] = make_float2(ss.hi, ss.lo);
    }
}

__global__ void k_meta_clear(ErosionMeta *meta)
{
    meta->it = 0;
    meta->load_max = 0;
    for (int k = 0; k < 3; ++k) {
        meta->E[k] = 0;
        meta->big[k] = 0;
    }
}

__global__ __launch_bounds__(256) void k_store(const ErosionMeta *__restrict__ meta,
                                               const float4 *__restrict__ hw0,
                                               const float4 *__restrict__ hw1,
                                               const float2 *__restrict__ s0,
                                               const float2 *__restrict__ s1, int64_t V,
                                               double *__restrict__ h, double *__restrict__ w,
                                               double *__restrict__ s)
{
    const int64_t it = meta->it;
    const int E = meta->E[(int)((it + 2) % 3)];
    const float4 *hw = (it & 1) ? hw1 : hw0;
    const float2 *sd = (it & 1) ? s1 : s0;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) {
        const float4 v = hw[i];
        if (h) h[i] = ldexp((double)v.x + (double)v.y, E);
        if (w) w[i] = ldexp((double)v.z + (double)v.w, E);
        if (s) {
            const float2 q = sd[i];
            s[i] = ldexp((double)q.x + (double)q.y, E);
        }
    }
}

// FP64 sums for the mass check, one partial per block then atomics.
__global__ __launch_bounds__(256) void k_mass(const ErosionMeta *__restrict__ meta,
                                              const float4 *__restrict__ hw0,
                                              const float4 *__restrict__ hw1,
                                              const float2 *__restrict__ s0,
                                              const float2 *__restrict__ s1, int64_t V,
                                              double *__restrict__ out)
{
    const int64_t it = meta->it;
    const int E = meta->E[(int)((it + 2) % 3)];
    const float4 *hw = (it & 1) ? hw1 : hw0;
    const float2 *sd = (it & 1) ? s1 : s0;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float4 v = hw[i];
        const float2 q = sd[i];
        const double hh = ldexp((double)v.x + (double)v.y, E);
        const double ss = ldexp((double)q.x + (double)q.y, E);
        a += hh + ss;
        b += fabs(hh) + fabs(ss);
        c += ldexp((double)v.z + (double)v.w, E);
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    __shared__ double red[3][4];
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
        red[2][threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[threadIdx.x][k];
        atomicAdd(&out[threadIdx.x], t);
    }
}

__global__ void k_zero3(double *out)
{
    if (threadIdx.x < 3) out[threadIdx.x] = 0.0;
}

// ---- host side -------------------------------------------------------------

static int check_state(nx_ctx *ctx, const void *state, int64_t V)
{
    if (!ctx) return set_error(NX_EINVAL, "null context");
    if (!state) return set_error(NX_EINVAL, "null erosion state");
    if (V <= 0 || V > INT32_MAX) return set_error(NX_EINVAL, "V=%lld outside 1..2^31-1", (long long)V);
    if (((uintptr_t)state & 255) != 0) return set_error(NX_EINVAL, "erosion state must be 256-byte aligned");
    return NX_OK;
}

int erosion_load(nx_ctx *ctx, void *state, int64_t V, const double *h, const double *w,
                 const double *s, hipStream_t stream)
{
    int rc = check_state(ctx, state, V);
    if (rc) return rc;
    if (!h) return set_error(NX_EINVAL, "null heights");
    NX_HIP(hipSetDevice(ctx->device));
    ErosionBufs b = bufs(state, V);
    const unsigned grid = (unsigned)((V + 255) / 256);
    hipLaunchKernelGGL(k_meta_clear, dim3(1), dim3(1), 0, stream, b.meta);
    hipLaunchKernelGGL(k_load_max, dim3(std::min<unsigned>(grid, 2048u)), dim3(256), 0, stream,
                       b.meta, h, w, s, V);
    hipLaunchKernelGGL(k_load, dim3(grid), dim3(256), 0, stream, b.meta, b.hw[0], b.s[0], h, w,
                       s, V);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

// Launch n consecutive steps (launch k uses step it + k, where `it` is read
// from the workspace) and advance the workspace's step counter by n.  The
// last step's rain for a following step is `rain` unless `last`.
static void launch_steps(const ErosionBufs &b, int64_t V, const int32_t *col, const float *elen,
                         int n, double rain, bool last, hipStream_t stream)
{
    const int64_t nblocks = (V + 255) / 256;
    const int64_t grid = ((nblocks + 7) / 8) * 8;
    for (int k = 0; k < n; ++k) {
        const double rain_next = (last && k + 1 == n) ? 0.0 : rain;
        hipLaunchKernelGGL(k_erode_step, dim3((unsigned)grid), dim3(256), 0, stream, b.meta,
                           b.hw[0], b.hw[1], b.s[0], b.s[1], V, col, (const float2 *)elen,
                           nblocks, (int64_t)k, rain_next);
    }
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, stream, b.meta, (int64_t)n);
}

// Captured chunk of steps, cached per (workspace, graph, length, rain, last).
// Replays stay correct because every launch reads the step counter, buffer
// parity and exponent state from the workspace rather than from its
// arguments.  No memset is recorded into the graph.
static int chunk_graph(nx_ctx *ctx, const ErosionBufs &b, void *state, int64_t V,
                       const int32_t *col, const float *elen, int n, double rain, bool last,
                       hipGraphExec_t *out)
{
    GraphKey key{state, V, col, elen, n, rain, last, 0};
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->graphs.find(key);
    if (it != ctx->graphs.end()) {
        *out = it->second;
        return NX_OK;
    }
    if (ctx->graphs.size() >= kMaxGraphs) ctx->clear_graphs();
    hipGraph_t g = nullptr;
    NX_HIP(hipStreamBeginCapture(ctx->capture_stream, hipStreamCaptureModeThreadLocal));
    launch_steps(b, V, col, elen, n, rain, last, ctx->capture_stream);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(ctx->capture_stream, &g);
    if (le != hipSuccess || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return set_error(NX_EHIP, "erosion graph capture failed: %s",
                         hipGetErrorString(le != hipSuccess ? le : ce));
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) return set_error(NX_EHIP, "graph instantiate: %s", hipGetErrorString(ie));
    ctx->graphs.emplace(key, ex);
    *out = ex;
    return NX_OK;
}

int erosion_run(nx_ctx *ctx, void *state, int64_t V, const int32_t *col, const float *elen,
                int iters, double rain, hipStream_t stream)
{
    int rc = check_state(ctx, state, V);
    if (rc) return rc;
    if (!col || !elen) return set_error(NX_EINVAL, "null neighbour table");
    if (iters < 0) return set_error(NX_EINVAL, "iters=%d < 0", iters);
    if (!(rain >= 0.0) || !std::isfinite(rain)) return set_error(NX_EINVAL, "rain must be finite and >= 0");
    if (iters == 0) return NX_OK;
    NX_HIP(hipSetDevice(ctx->device));
    ErosionBufs b = bufs(state, V);
    const int64_t nblocks = (V + 255) / 256;
    if (rain != 0.0)
        hipLaunchKernelGGL(k_add_rain, dim3((unsigned)nblocks), dim3(256), 0, stream, b.meta,
                           b.hw[0], b.hw[1], V, rain);
    const int chunk = ctx->graph_chunk;
    if (chunk <= 0 || iters < 2) {
        launch_steps(b, V, col, elen, iters, rain, true, stream);
        NX_HIP(hipGetLastError());
        return NX_OK;
    }
    // Full chunks, then the remainder; the final piece does not pre-add rain.
    int done = 0;
    while (done < iters) {
        const int n = std::min(chunk, iters - done);
        const bool last = done + n == iters;
        hipGraphExec_t ex = nullptr;
        rc = chunk_graph(ctx, b, state, V, col, elen, n, rain, last, &ex);
        if (rc) return rc;
        NX_HIP(hipGraphLaunch(ex, stream));
        done += n;
    }
    return NX_OK;
}

}  // namespace nx

namespace nx {

int erosion_store(nx_ctx *ctx, const void *state, int64_t V, double *h, double *w, double *s,
                  hipStream_t stream)
{
    int rc = check_state(ctx, state, V);
    if (rc) return rc;
    NX_HIP(hipSetDevice(ctx->device));
    ErosionBufs b = bufs(const_cast<void *>(state), V);
    hipLaunchKernelGGL(k_store, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, stream, b.meta,
                       b.hw[0], b.hw[1], b.s[0], b.s[1], V, h, w, s);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

int erosion_mass(nx_ctx *ctx, const void *state, int64_t V, double *out, hipStream_t stream)
{
    int rc = check_state(ctx, state, V);
    if (rc) return rc;
    if (!out) return set_error(NX_EINVAL, "null output");
    NX_HIP(hipSetDevice(ctx->device));
    ErosionBufs b = bufs(const_cast<void *>(state), V);
    hipLaunchKernelGGL(k_zero3, dim3(1), dim3(64), 0, stream, out);
    const unsigned grid = (unsigned)std::min<int64_t>((V + 255) / 256, 1024);
    hipLaunchKernelGGL(k_mass, dim3(grid), dim3(256), 0, stream, b.meta, b.hw[0], b.hw[1], b.s[0],
                       b.s[1], V, out);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

}  // namespace nx
