// nx_erosion.hip -- erosion Jacobi steps over the icosphere neighbour table.
//
// erosion.py:171-192 (erode_terrain3) adds rain to every vertex and runs
// erosion.py:197-279 (erosion_iteration3) once per iteration; the latter
// copies h/w/s, gathers every vertex's neighbours, and copies the results
// back.  Here the fields stay on the device in a double-buffered workspace:
//   * FP64 fields in the reference's operation order (nx_erosion_dev.h):
//     height+water packed as double2 (one 16-byte gather per neighbour),
//     sediment as double;
//   * neighbours come from the slot-major table col[6][V] with FP64 edge
//     lengths elen[6][V] (coalesced per slot);
//   * k_erode_step runs one step per launch; with a tile plan (nx_tiles.cpp)
//     k_erode_pair runs two steps per launch over LDS-resident tiles, which
//     halves the HBM traffic per step;
//   * the next iteration's rain is folded into this step's water output.
// Workspace: [meta 256 B][hw 2 x V x double2][s 2 x V x double].
#include <hip/hip_runtime.h>

#include "nx_ctx.h"
#include "nx_erosion_dev.h"
#include "nx_pair_dev.h"

namespace nx {

struct ErosionBufs {
    ErosionMeta *meta;
    double2 *hw[2];
    double *s[2];
};

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t erosion_state_bytes(int64_t V)
{
    if (V <= 0) return 256;
    return 256 + 2 * align256((size_t)V * sizeof(double2)) + 2 * align256((size_t)V * sizeof(double));
}

static ErosionBufs bufs(void *state, int64_t V)
{
    char *p = (char *)state;
    ErosionBufs b;
    b.meta = (ErosionMeta *)p;
    p += 256;
    for (int k = 0; k < 2; ++k) {
        b.hw[k] = (double2 *)p;
        p += align256((size_t)V * sizeof(double2));
    }
    for (int k = 0; k < 2; ++k) {
        b.s[k] = (double *)p;
        p += align256((size_t)V * sizeof(double));
    }
    return b;
}

// Launch j of a sequence reads the buffer that j earlier launches flipped
// to; the step counter and buffer come from the workspace so that replays
// of a captured sequence stay correct.  RULE 3 or 4 (nx_erosion_dev.h).
template <int RULE>
__global__ __launch_bounds__(256) void k_erode_step(const ErosionMeta *__restrict__ meta,
                                                    double2 *__restrict__ hw0,
                                                    double2 *__restrict__ hw1,
                                                    double *__restrict__ s0,
                                                    double *__restrict__ s1, int64_t V,
                                                    const int32_t *__restrict__ col,
                                                    const double *__restrict__ elen,
                                                    int64_t nblocks, int j, double rain_next)
{
    const int p = (meta->buf + j) & 1;
    const int64_t tile = xcd_tile(blockIdx.x, nblocks);
    const int64_t i = tile * blockDim.x + threadIdx.x;
    if (tile < nblocks && i < V)
        erode_vertex_step<RULE>(i, V, p ? hw1 : hw0, p ? s1 : s0, p ? hw0 : hw1, p ? s0 : s1, col,
                                elen, rain_next);
}

// Two steps per launch, one block per tile of the plan (nx_pair_dev.h).
// B holds step one's results, at most kStepOneMax (tile + ring 1) vertices.
template <int THREADS, bool KEEP, int RULE = 3>
__global__ __launch_bounds__(THREADS) void k_erode_pair(const ErosionMeta *__restrict__ meta,
                                                       double2 *__restrict__ hw0,
                                                       double2 *__restrict__ hw1,
                                                       double *__restrict__ s0,
                                                       double *__restrict__ s1, TilePlanDev P,
                                                       int j, double rain_mid, double rain_next)
{
    __shared__ double2 A[kLoadMax];
    __shared__ double2 B[THREADS == 1024 ? kStepOneMax : 1];
    const int64_t tile = xcd_tile(blockIdx.x, P.ntiles);
    if (tile >= P.ntiles) return;
    erode_tile_pair<THREADS, KEEP, RULE>(A, THREADS == 1024 ? B : nullptr, hw0, hw1, s0, s1,
                                         (meta->buf + j) & 1, P, tile, rain_mid, rain_next);
}

// Kernel variants (NX_OPT_PAIR_KERNEL): block size and register use.
struct PairVariant {
    void (*fn)(const ErosionMeta *, double2 *, double2 *, double *, double *, TilePlanDev, int,
               double, double);
    int threads;
};
static const PairVariant kPairVariants[] = {
    {k_erode_pair<1024, true>, 1024}, {k_erode_pair<512, true>, 512},
    {k_erode_pair<512, false>, 512},  {k_erode_pair<1024, true>, 1024},
    {k_erode_pair<1024, false>, 1024}, {k_erode_pair<256, true>, 256},
    {k_erode_pair<256, false>, 256},
};
// erosion_iteration4 uses the default variant only.
static const PairVariant kPairRule4 = {k_erode_pair<1024, true, 4>, 1024};

// kMultiSteps steps per launch over a plan built for them (nx_pair_dev.h).
template <int THREADS, bool KEEP, int RULE = 3>
__global__ __launch_bounds__(THREADS) void k_erode_multi(const ErosionMeta *__restrict__ meta,
                                                        double2 *__restrict__ hw0,
                                                        double2 *__restrict__ hw1,
                                                        double *__restrict__ s0,
                                                        double *__restrict__ s1, TilePlanDev P,
                                                        int j, double rain_mid, double rain_next)
{
    __shared__ double2 A[kLoadMaxMulti];
    const int64_t tile = xcd_tile(blockIdx.x, P.ntiles);
    if (tile >= P.ntiles) return;
    erode_tile_multi<THREADS, kMultiSteps, KEEP, RULE>(A, hw0, hw1, s0, s1, (meta->buf + j) & 1, P,
                                                      tile, rain_mid, rain_next);
}

// Four steps per launch with each thread's first-step data in registers and
// the steps alternating between two LDS buffers (nx_pair_dev.h); the plan
// keeps tile + rings 1..3 within kQuadStepOne and tile + rings within
// kQuadLoadMax.  NS = 2 or 3: the last two or three of those steps.
template <int RULE = 3, int NS = 4>
__global__ __launch_bounds__(1024) void k_erode_quad(const ErosionMeta *__restrict__ meta,
                                                    double2 *__restrict__ hw0,
                                                    double2 *__restrict__ hw1,
                                                    double *__restrict__ s0,
                                                    double *__restrict__ s1, TilePlanDev P, int j,
                                                    double rain_mid, double rain_next)
{
    __shared__ double2 A[kQuadLoadMax];
    __shared__ double2 B[kQuadStepOne];
    const int64_t tile = xcd_tile(blockIdx.x, P.ntiles);
    if (tile >= P.ntiles) return;
    erode_tile_quad<RULE, NS>(A, B, hw0, hw1, s0, s1, (meta->buf + j) & 1, P, tile, rain_mid,
                              rain_next);
}

// NX_OPT_MULTI_KERNEL variants: 0 keeps each thread's first-step data in
// registers (the default); 1..4 stage the slot tables in LDS and keep
// neighbour data in registers (KEEP) or not, in 1024- or 512-thread blocks.
// erosion_iteration4 uses the first.  A run's last two or three steps on a
// four-step plan take the tail launches.
static const PairVariant kMultiVariants[] = {
    {k_erode_quad<3>, 1024},
    {k_erode_multi<1024, false>, 1024}, {k_erode_multi<1024, true>, 1024},
    {k_erode_multi<512, false>, 512},   {k_erode_multi<512, true>, 512},
};
static const PairVariant kMultiRule4 = {k_erode_quad<4>, 1024};
static const PairVariant kQuadTails[2][2] = {   // [rule 4][steps 3]
    {{k_erode_quad<3, 2>, 1024}, {k_erode_quad<3, 3>, 1024}},
    {{k_erode_quad<4, 2>, 1024}, {k_erode_quad<4, 3>, 1024}},
};

// erosion.py:76-99 (erosion_iteration1): w[i] = r[i] + amt with amt moved
// by 0.0005 towards each higher neighbour and away from each lower one, in
// slot order (a pad slot points at the vertex itself and changes nothing,
// as the reference skipping -1).
__global__ __launch_bounds__(256) void k_erode_v1(const double *__restrict__ r,
                                                  double *__restrict__ w, int64_t V,
                                                  const int32_t *__restrict__ col,
                                                  int64_t nblocks)
{
    const int64_t tile = xcd_tile(blockIdx.x, nblocks);
    const int64_t i = tile * blockDim.x + threadIdx.x;
    if (tile >= nblocks || i >= V) return;
    const double simple_constant = 0.0005;
    const double thisvert = r[i];
    double amt = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double rn = r[col[(int64_t)k * V + i]];
        if (rn > thisvert)
            amt += simple_constant;
        else if (rn < thisvert)
            amt -= simple_constant;
    }
    w[i] = thisvert + amt;
}


// erosion.py:184-186: water += rain on the current fields (before a step).
__global__ __launch_bounds__(256) void k_add_rain(const ErosionMeta *__restrict__ meta,
                                                  double2 *__restrict__ hw0,
                                                  double2 *__restrict__ hw1, int64_t V,
                                                  double rain)
{
    double2 *hw = meta->buf ? hw1 : hw0;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) hw[i].y = hw[i].y + rain;
}

// After a sequence of launches: `steps` steps taken, the buffers flipped
// once per launch (one thread).
__global__ void k_advance(ErosionMeta *meta, int64_t steps, int launches)
{
    meta->it += steps;
    meta->buf ^= launches & 1;
}

// ---- load / store --------------------------------------------------------

__global__ __launch_bounds__(256) void k_load(ErosionMeta *__restrict__ meta,
                                              double2 *__restrict__ hw, double *__restrict__ sd,
                                              const double *__restrict__ h,
                                              const double *__restrict__ w,
                                              const double *__restrict__ s, int64_t V)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        meta->it = 0;
        meta->buf = 0;
    }
    if (i < V) {
        hw[i] = make_double2(h[i], w ? w[i] : 0.0);
        sd[i] = s ? s[i] : 0.0;
    }
}

__global__ __launch_bounds__(256) void k_store(const ErosionMeta *__restrict__ meta,
                                               const double2 *__restrict__ hw0,
                                               const double2 *__restrict__ hw1,
                                               const double *__restrict__ s0,
                                               const double *__restrict__ s1, int64_t V,
                                               double *__restrict__ h, double *__restrict__ w,
                                               double *__restrict__ s)
{
    const double2 *hw = meta->buf ? hw1 : hw0;
    const double *sd = meta->buf ? s1 : s0;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) {
        const double2 v = hw[i];
        if (h) h[i] = v.x;
        if (w) w[i] = v.y;
        if (s) s[i] = sd[i];
    }
}

// FP64 sums for the mass check, one partial per block then atomics.
__global__ __launch_bounds__(256) void k_mass(const ErosionMeta *__restrict__ meta,
                                              const double2 *__restrict__ hw0,
                                              const double2 *__restrict__ hw1,
                                              const double *__restrict__ s0,
                                              const double *__restrict__ s1, int64_t V,
                                              double *__restrict__ out)
{
    const double2 *hw = meta->buf ? hw1 : hw0;
    const double *sd = meta->buf ? s1 : s0;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = hw[i];
        const double ss = sd[i];
        a += v.x + ss;
        b += fabs(v.x) + fabs(ss);
        c += v.y;
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
    hipLaunchKernelGGL(k_load, dim3(grid), dim3(256), 0, stream, b.meta, b.hw[0], b.s[0], h, w,
                       s, V);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

// Launch n consecutive steps (the launches read the step counter and the
// current buffer from the workspace) and advance the workspace by n steps.
// With a two-step plan, steps go two per launch and an odd last step runs
// alone; with a four-step plan, four per launch and a last two or three in
// one tail launch.  The last step's rain for a following step is `rain`
// unless `last`.
static void launch_steps(const ErosionBufs &b, int64_t V, const int32_t *col, const double *elen,
                         const nx_tile_plan *plan, int variant, int mvariant, int rule, int n,
                         double rain, bool last, hipStream_t stream)
{
    const int64_t nblocks = (V + 255) / 256;
    const int64_t grid = ((nblocks + 7) / 8) * 8;
    int k = 0, j = 0;
    if (plan && plan->steps == kMultiSteps) {
        const TilePlanDev P = plan_dev(*plan);
        const int64_t tgrid = ((P.ntiles + 7) / 8) * 8;
        const PairVariant &mv = rule == 4 ? kMultiRule4 : kMultiVariants[mvariant];
        for (; k + kMultiSteps <= n; k += kMultiSteps, ++j) {
            const double rain_next = (last && k + kMultiSteps == n) ? 0.0 : rain;
            hipLaunchKernelGGL(mv.fn, dim3((unsigned)tgrid), dim3(mv.threads), 0, stream, b.meta,
                               b.hw[0], b.hw[1], b.s[0], b.s[1], P, j, rain, rain_next);
        }
        if (n - k >= 2) {
            const PairVariant &tv = kQuadTails[rule == 4][n - k - 2];
            const double rain_next = last ? 0.0 : rain;
            hipLaunchKernelGGL(tv.fn, dim3((unsigned)tgrid), dim3(tv.threads), 0, stream, b.meta,
                               b.hw[0], b.hw[1], b.s[0], b.s[1], P, j, rain, rain_next);
            k = n;
            ++j;
        }
    } else if (plan) {
        const TilePlanDev P = plan_dev(*plan);
        const int64_t tgrid = ((P.ntiles + 7) / 8) * 8;
        for (; k + 2 <= n; k += 2, ++j) {
            const double rain_next = (last && k + 2 == n) ? 0.0 : rain;
            const PairVariant &pv = rule == 4 ? kPairRule4 : kPairVariants[variant];
            hipLaunchKernelGGL(pv.fn, dim3((unsigned)tgrid), dim3(pv.threads), 0, stream, b.meta,
                               b.hw[0], b.hw[1], b.s[0], b.s[1], P, j, rain, rain_next);
        }
    }
    for (; k < n; ++k, ++j) {
        const double rain_next = (last && k + 1 == n) ? 0.0 : rain;
        hipLaunchKernelGGL(rule == 4 ? k_erode_step<4> : k_erode_step<3>, dim3((unsigned)grid),
                           dim3(256), 0, stream, b.meta, b.hw[0], b.hw[1], b.s[0], b.s[1], V, col,
                           elen, nblocks, j, rain_next);
    }
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, stream, b.meta, (int64_t)n, j);
}

// Everything a captured launch takes from the plan by value (TilePlanDev:
// the blob address, the section offsets, the tile count) plus the kernel
// variant and rule, folded into one 64-bit tag (FNV-1a).  Two plans that
// differ in any of these get different graphs even when a later plan's
// image lands at a freed plan's address.
static uint64_t plan_tag(const nx_tile_plan *plan, int variant, int rule)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](uint64_t v) {
        for (int k = 0; k < 8; ++k) {
            h ^= (v >> (8 * k)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    mix((uint64_t)rule);
    mix((uint64_t)variant);
    if (plan) {
        mix((uint64_t)(uintptr_t)plan->blob);
        const int64_t f[] = {plan->bytes,   plan->V,        plan->ntiles,    plan->max_tile,
                             plan->max_load, plan->off_start, plan->off_cs,   plan->off_load,
                             plan->off_rbase, plan->off_ring, plan->off_cbase, plan->off_lcol,
                             plan->off_lel,  plan->steps,    plan->max_step1, plan->off_pref,
                             plan->off_ebase};
        for (int64_t v : f) mix((uint64_t)v);
    }
    return h;
}

// Captured chunk of steps, cached per (workspace, graph, plan contents,
// length, rain, last).  Replays stay correct because every launch reads the
// step counter and current buffer from the workspace rather than from its
// arguments.  No memset is recorded into the graph.
static int chunk_graph(nx_ctx *ctx, const ErosionBufs &b, void *state, int64_t V,
                       const int32_t *col, const double *elen, const nx_tile_plan *plan, int rule,
                       int n, double rain, bool last, hipGraphExec_t *out)
{
    const int variant =
        plan ? (plan->steps == kMultiSteps ? 8 + ctx->multi_variant : ctx->pair_variant) : -1;
    GraphKey key{state, V, col, elen, n, rain, last, plan_tag(plan, variant, rule)};
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->graphs.find(key);
    if (it != ctx->graphs.end()) {
        *out = it->second;
        return NX_OK;
    }
    ctx->make_room_for_graph();
    hipGraph_t g = nullptr;
    NX_HIP(hipStreamBeginCapture(ctx->capture_stream, hipStreamCaptureModeThreadLocal));
    launch_steps(b, V, col, elen, plan, ctx->pair_variant, ctx->multi_variant, rule, n, rain, last,
                 ctx->capture_stream);
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

static int check_plan(const nx_tile_plan *plan, int64_t V)
{
    if (!plan) return NX_OK;
    if (!plan->blob || ((uintptr_t)plan->blob & 255) != 0)
        return set_error(NX_EINVAL, "tile plan needs a 256-byte aligned device copy");
    if (plan->V != V || plan->ntiles <= 0 || plan->ntiles > INT32_MAX)
        return set_error(NX_EINVAL, "tile plan is for V=%lld, not %lld", (long long)plan->V,
                         (long long)V);
    if (plan->steps != 2 && plan->steps != kMultiSteps)
        return set_error(NX_EINVAL, "tile plan for %d steps per launch (2 or %d)", plan->steps,
                         kMultiSteps);
    const int lmax = plan->steps == 2 ? kLoadMax : kLoadMaxMulti;
    const int tmax = plan->steps == 2 ? kTileMax : kQuadStepOne;
    if (plan->max_tile > tmax || plan->max_load > lmax)
        return set_error(NX_EINVAL, "tile plan tiles exceed %d owned / %d loaded vertices", tmax,
                         lmax);
    const int s1max = plan->steps == 2 ? kStepOneMax : kQuadStepOne;
    if (plan->max_step1 > s1max || (plan->steps != 2 && plan->max_load > kQuadLoadMax))
        return set_error(NX_EINVAL, "tile plan tiles hold %d first-step and %d loaded vertices "
                         "(at most %d and %d)", plan->max_step1, plan->max_load, s1max,
                         plan->steps == 2 ? kLoadMax : kQuadLoadMax);
    return NX_OK;
}

int erosion_run(nx_ctx *ctx, void *state, int64_t V, const int32_t *col, const double *elen,
                const nx_tile_plan *plan, int iters, double rain, hipStream_t stream)
{
    return erosion_run_rule(ctx, state, V, col, elen, plan, iters, rain, 3, stream);
}

int erosion_run_rule(nx_ctx *ctx, void *state, int64_t V, const int32_t *col, const double *elen,
                     const nx_tile_plan *plan, int iters, double rain, int rule,
                     hipStream_t stream)
{
    int rc = check_state(ctx, state, V);
    if (rc) return rc;
    if (rule != 3 && rule != 4) return set_error(NX_EINVAL, "erosion rule %d is not 3 or 4", rule);
    if (!col || !elen) return set_error(NX_EINVAL, "null neighbour table");
    if ((rc = check_plan(plan, V))) return rc;
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
        launch_steps(b, V, col, elen, plan, ctx->pair_variant, ctx->multi_variant, rule, iters, rain,
                     true, stream);
        NX_HIP(hipGetLastError());
        return NX_OK;
    }
    // Full chunks, then the remainder; the final piece does not pre-add rain.
    int done = 0;
    while (done < iters) {
        const int n = std::min(chunk, iters - done);
        const bool last = done + n == iters;
        hipGraphExec_t ex = nullptr;
        rc = chunk_graph(ctx, b, state, V, col, elen, plan, rule, n, rain, last, &ex);
        if (rc) return rc;
        NX_HIP(hipGraphLaunch(ex, stream));
        done += n;
    }
    return NX_OK;
}

}  // namespace nx

namespace nx {

// erosion.py:42-70 (erode_terrain1): iters steps of erosion_iteration1,
// Jacobi, in place on h, with tmp (V doubles) as the other buffer.
int erosion1_run(nx_ctx *ctx, double *h, double *tmp, int64_t V, const int32_t *col, int iters,
                 hipStream_t stream)
{
    if (!ctx || !h || !tmp || !col) return set_error(NX_EINVAL, "null argument");
    if (V <= 0 || V > INT32_MAX) return set_error(NX_EINVAL, "V=%lld outside 1..2^31-1", (long long)V);
    if (iters < 0) return set_error(NX_EINVAL, "iters=%d < 0", iters);
    if (iters == 0) return NX_OK;
    NX_HIP(hipSetDevice(ctx->device));
    const int64_t nblocks = (V + 255) / 256;
    const int64_t grid = ((nblocks + 7) / 8) * 8;
    double *src = h, *dst = tmp;
    for (int k = 0; k < iters; ++k) {
        hipLaunchKernelGGL(k_erode_v1, dim3((unsigned)grid), dim3(256), 0, stream, src, dst, V, col,
                           nblocks);
        std::swap(src, dst);
    }
    if (src != h) NX_HIP(hipMemcpyAsync(h, src, sizeof(double) * (size_t)V, hipMemcpyDeviceToDevice, stream));
    NX_HIP(hipGetLastError());
    return NX_OK;
}

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
