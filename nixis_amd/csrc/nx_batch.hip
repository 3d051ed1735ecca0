// nx_batch.hip -- many planets per launch.
//
// Batched form of the terrain path for jobs of uneven size, seed, octave
// count and erosion length:
//   * noise: one persistent launch for all planets.  Work items are
//     (planet, tile) pairs handed out by an atomic counter; planets are
//     ordered by octave count (cost per tile) so expensive tiles go first,
//     and a block reloads the seed tables only when it moves to another
//     planet.  Tile size and blocks per CU come from the autotuner.
//   * rescale + erosion load: per-planet min/max (one atomic per block per
//     planet) then the map of util.py:139, which also writes the planet's
//     FP64 erosion state with the first rain already added.
//   * erosion: a tile phase, then a one-step phase.  Launch k of the tile
//     phase covers every planet with more than k tile launches in this call
//     (planets ordered by that count, so the active set is a prefix): four
//     steps of a four-step plan (the last two or three in one tail launch)
//     or two of a two-step plan; a step left over runs in the one-step
//     phase.  Each planet keeps its own step counter and buffer index in
//     its workspace, so a planet's result does not depend on its batch.
//     The launch sequence can be captured into a hipGraph.
// Each planet's arithmetic is the single-planet kernels' arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <initializer_list>
#include <numeric>
#include <vector>

#include "nx_ctx.h"
#include "nx_erosion_dev.h"
#include "nx_pair_dev.h"
#include "nx_noise_dev.h"
#include "nx_reduce_dev.h"

namespace nx {

int pack_tables(const int32_t *perm, const int32_t *pgi, NoiseTablesPacked &p);
int octave_plan(const double *rough, const double *strength, double radius, int n_oct,
                OctavePlan &p);

struct BatchPlanetDev {
    const double *verts;
    double *elev;
    int64_t V;
    int32_t accumulate;
    double radius;
    double lower, upper;
    int32_t do_rescale;
    int32_t do_erode;
    ErosionMeta *meta;
    double2 *hw0, *hw1;
    double *s0, *s1;
    const int32_t *col;
    const double *elen;
    int64_t total_iters;
    double rain;
    int64_t nblocks;  // ceil(V / 256)
    TilePlanDev plan;
    int32_t has_plan;
    // this call's tile launches: full_launches of plan.steps steps, then
    // one of tail_steps (2 or 3; 0: none)
    int32_t tail_steps;
    int64_t full_launches;
    OctavePlan oct;
    NoiseTablesPacked tables;
};

constexpr int kRsTile = 4096;  // vertices per block in the rescale / load / store passes

// Workspace layout (offsets in bytes, all 256-aligned).
struct BatchLayout {
    size_t counter, keys, desc, noise_prefix, rs_prefix, tile_prefix[4], tile_order[4],
        tile_bucket[4], one_prefix, one_order, one_base, steps, launches, total;
};

// Tile phases of an erosion call, in launch order (a planet's tail runs
// before its four-step launches).
enum TilePhaseKind { kPhaseTail3, kPhaseTail2, kPhaseQuad, kPhasePair, kTilePhases };

// Planet lookup of the many-planet tile kernels: tile b falls in bucket
// b >> shift; the bucket holds the planet (in the phase's order) of its first tile
// and the first tile of the next planet.
constexpr int kPairBuckets = 1024;
struct PairBucket {
    int32_t r;
    int32_t next;
};

static inline size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

static BatchLayout layout(int n)
{
    BatchLayout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += a256(bytes);
        return at;
    };
    L.counter = take(256);
    L.keys = take(sizeof(unsigned long long) * 2 * n);
    L.desc = take(sizeof(BatchPlanetDev) * n);
    L.noise_prefix = take(sizeof(int64_t) * (n + 1));
    L.rs_prefix = take(sizeof(int64_t) * (n + 1));
    for (int f = 0; f < kTilePhases; ++f) {
        L.tile_prefix[f] = take(sizeof(int64_t) * (n + 1));
        L.tile_order[f] = take(sizeof(int32_t) * n);
        L.tile_bucket[f] = take(sizeof(PairBucket) * kPairBuckets);
    }
    L.one_prefix = take(sizeof(int64_t) * (n + 1));
    L.one_order = take(sizeof(int32_t) * n);
    L.one_base = take(sizeof(int64_t) * n);
    L.steps = take(sizeof(int64_t) * n);
    L.launches = take(sizeof(int32_t) * n);
    L.total = o;
    return L;
}

size_t batch_workspace_bytes(int n) { return n > 0 ? layout(n).total : 256; }

// Index r with prefix[r] <= x < prefix[r + 1] (prefix ascending, n entries).
__device__ __forceinline__ int find_slot(const int64_t *prefix, int n, int64_t x)
{
    int lo = 0, hi = n;  // invariant: prefix[lo] <= x < prefix[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ void k_batch_reset(unsigned *counter, unsigned long long *keys, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *counter = 0u;
    if (i < n) {
        keys[2 * i] = ~0ull;
        keys[2 * i + 1] = 0ull;
    }
}

__device__ __forceinline__ void load_seed_tables(NoiseTables &T, const NoiseTablesPacked &P)
{
    const uint8_t *pb = reinterpret_cast<const uint8_t *>(P.perm);
    const uint8_t *gb = reinterpret_cast<const uint8_t *>(P.grad);
    for (int i = threadIdx.x; i < kWrap; i += blockDim.x) {
        T.perm[i] = pb[(i + 255) & 255];
        T.grad[i] = gb[(i + 255) & 255];
    }
    uint32_t *hw = reinterpret_cast<uint32_t *>(T.g16);
    for (int i = threadIdx.x; i < 48; i += blockDim.x) hw[i] = P.g16[i];
}

// 8 waves per SIMD (at most 64 VGPRs), as for the single-planet noise launch.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_noise_batch(const BatchPlanetDev *__restrict__ P,
                                                     const int64_t *__restrict__ prefix, int np,
                                                     int64_t total_tiles, int tile,
                                                     unsigned *__restrict__ counter,
                                                     int *__restrict__ err)
{
    __shared__ NoiseTables T;
    __shared__ int64_t s_item;
    __shared__ int s_p;
    load_ext_table(T);
    int cur = -1;
    bool bad = false;
    for (;;) {
        if (threadIdx.x == 0) {
            const int64_t item = (int64_t)atomicAdd(counter, 1u);
            s_item = item;
            s_p = item < total_tiles ? find_slot(prefix, np, item) : -1;
        }
        __syncthreads();
        const int64_t item = s_item;
        const int p = s_p;
        if (item >= total_tiles) break;
        if (p != cur) {
            load_seed_tables(T, P[p].tables);
            __syncthreads();
            cur = p;
        }
        const BatchPlanetDev &D = P[p];
        const int64_t v0 = (item - prefix[p]) * tile;
        const int64_t v1 = v0 + tile < D.V ? v0 + tile : D.V;
        const auto verts = gmem(D.verts);
        const auto elev = gmem(D.elev);
        for (int64_t i = v0 + threadIdx.x; i < v1; i += blockDim.x) {
            const double x = verts[3 * i + 0], y = verts[3 * i + 1], z = verts[3 * i + 2];
            const double acc0 = D.accumulate ? elev[i] : 0.0;
            elev[i] = octave_sum<false>(x, y, z, acc0, D.oct, T, bad,
                                        [](int, float, const NoiseSel &) {});
        }
        __syncthreads();
    }
    if (bad) __hip_atomic_fetch_or(err, NX_ERR_RANGE_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-planet min/max of the noise output (planets that rescale or erode).
__global__ __launch_bounds__(256) void k_batch_minmax(const BatchPlanetDev *__restrict__ P,
                                                      const int64_t *__restrict__ prefix, int np,
                                                      unsigned long long *__restrict__ keys)
{
    const int p = find_slot(prefix, np, blockIdx.x);
    const BatchPlanetDev &D = P[p];
    if (!D.do_rescale && !D.do_erode) return;
    const int64_t v0 = (blockIdx.x - prefix[p]) * (int64_t)kRsTile;
    const int64_t v1 = v0 + kRsTile < D.V ? v0 + kRsTile : D.V;
    unsigned long long lo = ~0ull, hi = 0ull;
    for (int64_t i = v0 + threadIdx.x; i < v1; i += blockDim.x) {
        const unsigned long long k = dkey(D.elev[i]);
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    }
    block_minmax(lo, hi);
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_min(&keys[2 * p], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_max(&keys[2 * p + 1], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Rescale (util.py:139) and load the erosion state (exponent fixed by the
// rescale bounds, or by the max |height| without rescale; first rain added).
__global__ __launch_bounds__(256) void k_batch_map(const BatchPlanetDev *__restrict__ P,
                                                   const int64_t *__restrict__ prefix, int np,
                                                   const unsigned long long *__restrict__ keys,
                                                   int *__restrict__ err)
{
    const int p = find_slot(prefix, np, blockIdx.x);
    const BatchPlanetDev &D = P[p];
    if (!D.do_rescale && !D.do_erode) return;
    const double x_min = dunkey(keys[2 * p]), x_max = dunkey(keys[2 * p + 1]);
    const double x_range = x_max - x_min, new_range = D.upper - D.lower;
    const int64_t v0 = (blockIdx.x - prefix[p]) * (int64_t)kRsTile;
    if (v0 == 0 && threadIdx.x == 0) {
        if (D.do_rescale && x_range == 0.0)
            __hip_atomic_fetch_or(err, NX_ERR_DIV0_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (D.do_erode) {
            D.meta->it = 0;
            D.meta->buf = 0;
        }
    }
    const int64_t v1 = v0 + kRsTile < D.V ? v0 + kRsTile : D.V;
    const double rain = D.total_iters > 0 ? D.rain : 0.0;
    for (int64_t i = v0 + threadIdx.x; i < v1; i += blockDim.x) {
        double h = D.elev[i];
        if (D.do_rescale) {
            h = ((h - x_min) / x_range) * new_range + D.lower;
            D.elev[i] = h;
        }
        if (D.do_erode) {
            // erosion.py:178-185: zero water and sediment, then the first rain
            D.hw0[i] = make_double2(h, 0.0 + rain);
            D.s0[i] = 0.0;
        }
    }
}

// One erosion step for the first `active` planets of `order`: launch k of
// the one-step phase, after base[r] tile launches of planet order[r].
__global__ __launch_bounds__(256) void k_erode_batch(const BatchPlanetDev *__restrict__ P,
                                                     const int32_t *__restrict__ order,
                                                     const int64_t *__restrict__ prefix,
                                                     const int64_t *__restrict__ base, int active,
                                                     int64_t nb_total, int64_t k)
{
    const int64_t b = xcd_tile(blockIdx.x, nb_total);
    if (b >= nb_total) return;
    const int r = find_slot(prefix, active, b);
    const BatchPlanetDev &D = P[order[r]];
    const ErosionMeta *meta = D.meta;
    const int64_t launch = base[r] + k;
    const int64_t t = meta->it + D.full_launches * D.plan.steps + D.tail_steps + k;
    const int p = (meta->buf + (int)(launch & 1)) & 1;
    const double2 *hw_src = p ? D.hw1 : D.hw0;
    const double *s_src = p ? D.s1 : D.s0;
    double2 *hw_dst = p ? D.hw0 : D.hw1;
    double *s_dst = p ? D.s0 : D.s1;
    const double rain = t + 1 < D.total_iters ? D.rain : 0.0;
    const int64_t i = (b - prefix[r]) * blockDim.x + threadIdx.x;
    if (i < D.V) erode_vertex_step(i, D.V, hw_src, s_src, hw_dst, s_dst, D.col, D.elen, rain);
}

// The tile phases.  A planet with a four-step plan takes, in this order,
// one tail launch of 2 or 3 steps (when its step count leaves one) and then
// its four-step launches; a planet with a two-step plan takes its two-step
// launches; a step left over runs in the one-step phase.  Each phase's
// launches run one tile kernel, a template on the plan's steps per launch K
// and the launch's steps NS: (4, 3) and (4, 2) are the tails, (4, 4) the
// four-step phase, (2, 2) the two-step phase.  Launch k of a phase covers
// every tile of the first `active` planets in the phase's order (planets
// ordered by their launch count in the phase, one block per tile) and runs
// steps t .. t + NS - 1, t = it + (tail before it) + NS k, reading buffer
// (buf + launches before it) & 1.
template <int K, int NS>
__device__ __forceinline__ void batch_tile(double2 *A, double2 *B, double2 *hw0, double2 *hw1,
                                           double *s0, double *s1, int64_t it, int buf,
                                           const TilePlanDev &plan, int tail, int64_t total_iters,
                                           double rain, int64_t tile, int64_t k)
{
    constexpr bool after_tail = K == 4 && NS == 4;
    const int64_t t = it + (after_tail ? tail : 0) + NS * k;
    const int64_t launch = (after_tail && tail ? 1 : 0) + k;
    const int p = (buf + (int)(launch & 1)) & 1;
    const double rain_mid = t + 1 < total_iters ? rain : 0.0;
    const double rain_next = t + NS < total_iters ? rain : 0.0;
    if constexpr (K == 2)
        erode_tile_pair<1024, true>(A, B, hw0, hw1, s0, s1, p, plan, tile, rain_mid, rain_next);
    else
        erode_tile_quad<3, NS>(A, B, hw0, hw1, s0, s1, p, plan, tile, rain_mid, rain_next);
}

template <int K>
struct TileLds {
    static constexpr int kA = K == 2 ? kLoadMax : kQuadLoadMax;
    static constexpr int kB = K == 2 ? kStepOneMax : kQuadStepOne;
};

// A launch that covers one planet takes the single-planet kernel's
// arguments (buffers and plan by value) with the planet's step count, rain
// and tail; the step counter and buffer index come from one load of the
// planet's meta, as in the single-planet kernel.
template <int K, int NS>
__global__ __launch_bounds__(1024) void k_erode_batch_tile_one(const ErosionMeta *__restrict__ meta,
                                                               double2 *__restrict__ hw0,
                                                               double2 *__restrict__ hw1,
                                                               double *__restrict__ s0,
                                                               double *__restrict__ s1,
                                                               TilePlanDev plan, int tail, int64_t k,
                                                               int64_t total_iters, double rain)
{
    __shared__ double2 A[TileLds<K>::kA];
    __shared__ double2 B[TileLds<K>::kB];
    const int64_t tile = xcd_tile(blockIdx.x, plan.ntiles);
    if (tile >= plan.ntiles) return;
    const ErosionMeta m = *meta;
    batch_tile<K, NS>(A, B, hw0, hw1, s0, s1, m.it, m.buf, plan, tail, total_iters, rain, tile, k);
}

// A launch over at most kArgPlanets planets takes what each block needs in
// its kernel arguments (tile prefix, plan, buffers, step counter, rain), so
// a block finds its planet among scalar argument loads instead of a search
// over device arrays followed by the descriptor.  The prefix is int32 and
// padded past the launch's planets with values no tile index reaches, so the
// planet is a count of prefix entries <= the tile: fixed argument offsets,
// loaded together and counted with scalar arithmetic.
constexpr int kArgPlanets = 16;
struct TilePlanetArgs {
    TilePlanDev plan;
    double2 *hw0, *hw1;
    double *s0, *s1;
    const ErosionMeta *meta;
    int64_t total_iters;
    double rain;
    int32_t tail_steps;
};
struct TileLaunchArgs {
    int32_t prefix[kArgPlanets];   // prefix[q] = first tile of planet q (q < active)
    TilePlanetArgs pl[kArgPlanets];
};
static_assert(sizeof(TileLaunchArgs) + 16 <= 4096, "kernel arguments within 4 KiB");

template <int K, int NS>
__global__ __launch_bounds__(1024) void k_erode_batch_tile_args(const TileLaunchArgs a,
                                                                int32_t nt_total, int64_t k)
{
    __shared__ double2 A[TileLds<K>::kA];
    __shared__ double2 B[TileLds<K>::kB];
    const int b = (int)xcd_tile(blockIdx.x, nt_total);
    if (b >= nt_total) return;
    // b and the prefix entries lie in [0, 2^31): b - prefix[q] is negative,
    // its sign bit set, exactly when planet q starts after tile b.
    int after = 0;
#pragma unroll
    for (int q = 1; q < kArgPlanets; ++q) after += (int)(((uint32_t)b - (uint32_t)a.prefix[q]) >> 31);
    const int r = (kArgPlanets - 1) - after;
    const TilePlanetArgs &Q = a.pl[r];
    batch_tile<K, NS>(A, B, Q.hw0, Q.hw1, Q.s0, Q.s1, Q.meta->it, Q.meta->buf, Q.plan, Q.tail_steps,
                      Q.total_iters, Q.rain, b - a.prefix[r], k);
}

// A launch over more planets reads their descriptors (PD[order[r]] for
// the phase's planet r; prefix[r] = its first tile) and finds its planet in
// the buckets.
template <int K, int NS>
__global__ __launch_bounds__(1024) void k_erode_batch_tile(const BatchPlanetDev *__restrict__ PD,
                                                           const int32_t *__restrict__ order,
                                                           const PairBucket *__restrict__ bucket,
                                                           int shift,
                                                           const int64_t *__restrict__ prefix,
                                                           int active, int64_t nt_total,
                                                           int64_t k)
{
    __shared__ double2 A[TileLds<K>::kA];
    __shared__ double2 B[TileLds<K>::kB];
    const int64_t b = xcd_tile(blockIdx.x, nt_total);
    if (b >= nt_total) return;
    // The planet is the same for the whole block: keep its descriptor in
    // scalar registers (a local copy, so stores to the fields cannot be
    // taken to alias it).  Its bucket gives it unless the bucket holds the
    // start of another planet at or before b.
    const PairBucket q = bucket[b >> shift];
    int r = q.r;
    if (b >= q.next) {
        ++r;
        while (r + 1 < active && prefix[r + 1] <= b) ++r;
    }
    r = __builtin_amdgcn_readfirstlane(r);
    const BatchPlanetDev &D = PD[order[r]];
    const TilePlanDev plan = D.plan;
    double2 *hw0 = D.hw0, *hw1 = D.hw1;
    double *s0 = D.s0, *s1 = D.s1;
    batch_tile<K, NS>(A, B, hw0, hw1, s0, s1, D.meta->it, D.meta->buf, plan, D.tail_steps,
                      D.total_iters, D.rain, b - prefix[r], k);
}

// The three launch forms of one tile phase.
template <int K, int NS>
struct TileKernels {
    decltype(&k_erode_batch_tile_one<K, NS>) one = k_erode_batch_tile_one<K, NS>;
    decltype(&k_erode_batch_tile_args<K, NS>) args = k_erode_batch_tile_args<K, NS>;
    decltype(&k_erode_batch_tile<K, NS>) many = k_erode_batch_tile<K, NS>;
};

// After both phases: planet p took steps[p] steps in launches[p] launches.
__global__ void k_batch_advance(const BatchPlanetDev *__restrict__ P,
                                const int64_t *__restrict__ steps,
                                const int32_t *__restrict__ launches, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n && P[q].do_erode && steps[q] > 0) {
        ErosionMeta *m = P[q].meta;
        m->it += steps[q];
        m->buf ^= launches[q] & 1;
    }
}

__global__ __launch_bounds__(256) void k_batch_store(const BatchPlanetDev *__restrict__ P,
                                                     const int64_t *__restrict__ prefix, int np)
{
    const int p = find_slot(prefix, np, blockIdx.x);
    const BatchPlanetDev &D = P[p];
    if (!D.do_erode) return;
    const double2 *hw = D.meta->buf ? D.hw1 : D.hw0;
    const int64_t v0 = (blockIdx.x - prefix[p]) * (int64_t)kRsTile;
    const int64_t v1 = v0 + kRsTile < D.V ? v0 + kRsTile : D.V;
    for (int64_t i = v0 + threadIdx.x; i < v1; i += blockDim.x) {
        D.elev[i] = hw[i].x;
    }
}

// ---- host side ---------------------------------------------------------

static int fill_desc(const nx_planet &q, BatchPlanetDev &d)
{
    std::memset(&d, 0, sizeof(d));
    if (q.V <= 0 || q.V > INT32_MAX || !q.verts || !q.elev)
        return set_error(NX_EINVAL, "planet needs 0 < V < 2^31 and device verts/elev");
    if (q.n_octaves < 0 || q.n_octaves > NX_MAX_OCTAVES || (q.n_octaves > 0 && (!q.rough || !q.strength)))
        return set_error(NX_EINVAL, "planet octaves %d outside 0..%d", q.n_octaves, NX_MAX_OCTAVES);
    int rc = pack_tables(q.perm, q.pgi, d.tables);
    if (rc) return rc;
    d.verts = q.verts;
    d.elev = q.elev;
    d.V = q.V;
    d.accumulate = q.accumulate ? 1 : 0;
    d.radius = q.radius;
    rc = octave_plan(q.rough, q.strength, q.radius, q.n_octaves, d.oct);
    if (rc) return rc;
    d.lower = q.lower;
    d.upper = q.upper;
    d.do_rescale = q.lower < q.upper;
    d.nblocks = (q.V + 255) / 256;
    if (q.erosion_state) {
        if (((uintptr_t)q.erosion_state & 255) || !q.col || !q.elen || q.erosion_iters < 0 ||
            !(q.rain >= 0.0))
            return set_error(NX_EINVAL, "planet erosion state must be 256-aligned with col/elen, "
                                        "iters >= 0 and rain >= 0");
        char *base = (char *)q.erosion_state;
        const size_t hw = (size_t)a256((size_t)q.V * sizeof(double2)), sd = a256((size_t)q.V * 8);
        d.do_erode = 1;
        d.meta = (ErosionMeta *)base;
        d.hw0 = (double2 *)(base + 256);
        d.hw1 = (double2 *)(base + 256 + hw);
        d.s0 = (double *)(base + 256 + 2 * hw);
        d.s1 = (double *)(base + 256 + 2 * hw + sd);
        d.col = q.col;
        d.elen = q.elen;
        d.total_iters = q.erosion_iters;
        d.rain = q.rain;
        if (q.plan) {
            const nx_tile_plan &pl = *q.plan;
            const bool two = pl.steps == 2;
            if (!pl.blob || ((uintptr_t)pl.blob & 255) || pl.V != q.V || pl.ntiles <= 0 ||
                pl.ntiles > INT32_MAX || (!two && pl.steps != kMultiSteps) ||
                pl.max_tile > (two ? kTileMax : kQuadStepOne) ||
                pl.max_load > (two ? kLoadMax : kQuadLoadMax) ||
                pl.max_step1 > (two ? kStepOneMax : kQuadStepOne))
                return set_error(NX_EINVAL, "planet tile plan is not a 2- or 4-step plan for V=%lld within the kernel limits",
                                 (long long)q.V);
            d.plan = plan_dev(pl);
            d.has_plan = 1;
        }
    }
    return NX_OK;
}

// Host arrays into their places in the workspace region [L.desc, L.total),
// in one asynchronous copy through a pinned staging buffer.
struct Part {
    size_t off;
    const void *src;
    size_t bytes;
};

template <class Parts>
static int upload_parts(nx_ctx *ctx, char *w, const BatchLayout &L, const Parts &parts,
                        hipStream_t stream)
{
    const size_t bytes = L.total - L.desc;
    int slot = 0;
    char *h = stage_begin(ctx, bytes, &slot);
    if (!h) return set_error(NX_EHIP, "pinned staging buffer unavailable");
    for (const Part &p : parts) std::memcpy(h + (p.off - L.desc), p.src, p.bytes);
    return stage_commit(ctx, slot, w + L.desc, bytes, stream);
}

static int upload(nx_ctx *ctx, char *w, const BatchLayout &L, std::initializer_list<Part> parts,
                  hipStream_t stream)
{
    return upload_parts(ctx, w, L, parts, stream);
}

int batch_noise(nx_ctx *ctx, nx_planet *planets, int n, void *ws, int tile, int blocks_per_cu,
                hipStream_t stream)
{
    if (!ctx || !planets || n <= 0 || !ws) return set_error(NX_EINVAL, "batch needs planets and a workspace");
    if (((uintptr_t)ws & 255) != 0) return set_error(NX_EINVAL, "workspace must be 256-byte aligned");
    if (tile < 256 || tile > (1 << 22) || (tile & 255)) return set_error(NX_EINVAL, "tile %d: multiple of 256 in 256..4M", tile);
    if (blocks_per_cu < 1 || blocks_per_cu > 16) return set_error(NX_EINVAL, "blocks_per_cu %d outside 1..16", blocks_per_cu);
    const BatchLayout L = layout(n);
    // Noise order: most octaves (most work per tile) first.
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return planets[a].n_octaves > planets[b].n_octaves; });
    std::vector<BatchPlanetDev> desc(n);
    std::vector<int64_t> tpre(n + 1, 0), rpre(n + 1, 0);
    for (int r = 0; r < n; ++r) {
        int rc = fill_desc(planets[ord[r]], desc[r]);
        if (rc) return rc;
        tpre[r + 1] = tpre[r] + (desc[r].V + tile - 1) / tile;
        rpre[r + 1] = rpre[r] + (desc[r].V + kRsTile - 1) / kRsTile;
    }
    char *w = (char *)ws;
    NX_HIP(hipSetDevice(ctx->device));
    int rc = upload(ctx, w, L,
                    {{L.desc, desc.data(), sizeof(BatchPlanetDev) * n},
                     {L.noise_prefix, tpre.data(), sizeof(int64_t) * (n + 1)},
                     {L.rs_prefix, rpre.data(), sizeof(int64_t) * (n + 1)}},
                    stream);
    if (rc) return rc;
    auto *counter = (unsigned *)(w + L.counter);
    auto *keys = (unsigned long long *)(w + L.keys);
    const auto *D = (const BatchPlanetDev *)(w + L.desc);
    hipLaunchKernelGGL(k_batch_reset, dim3((n + 255) / 256), dim3(256), 0, stream, counter, keys, n);
    const int64_t grid = std::min<int64_t>(tpre[n], (int64_t)ctx->num_cus * blocks_per_cu);
    hipLaunchKernelGGL(k_noise_batch, dim3((unsigned)std::max<int64_t>(grid, 1)), dim3(256), 0, stream,
                       D, (const int64_t *)(w + L.noise_prefix), n, tpre[n], tile, counter, ctx->d_err);
    bool need_map = false;
    for (int r = 0; r < n; ++r) need_map = need_map || desc[r].do_rescale || desc[r].do_erode;
    if (need_map) {
        const int64_t *rp = (const int64_t *)(w + L.rs_prefix);
        hipLaunchKernelGGL(k_batch_minmax, dim3((unsigned)rpre[n]), dim3(256), 0, stream, D, rp, n, keys);
        hipLaunchKernelGGL(k_batch_map, dim3((unsigned)rpre[n]), dim3(256), 0, stream, D, rp, n,
                           (const unsigned long long *)keys, ctx->d_err);
    }
    NX_HIP(hipGetLastError());
    for (int p = 0; p < n; ++p) planets[p].steps_done = 0;
    return NX_OK;
}

static uint64_t fnv(const void *p, size_t n, uint64_t h)
{
    const unsigned char *c = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
    return h;
}

int batch_erode(nx_ctx *ctx, nx_planet *planets, int n, int max_steps, void *ws, int use_graph,
                hipStream_t stream)
{
    if (!ctx || !planets || n <= 0 || !ws) return set_error(NX_EINVAL, "batch needs planets and a workspace");
    if (((uintptr_t)ws & 255) != 0) return set_error(NX_EINVAL, "workspace must be 256-byte aligned");
    if (max_steps < 0) return set_error(NX_EINVAL, "max_steps < 0");
    const BatchLayout L = layout(n);
    std::vector<BatchPlanetDev> desc(n);
    for (int p = 0; p < n; ++p) {
        int rc = fill_desc(planets[p], desc[p]);
        if (rc) return rc;
    }
    // Steps of each planet in this call: a tail launch of 2 or 3 steps and
    // then four-step launches with a four-step plan, two-step launches with
    // a two-step plan, and a last step (or every step, without a plan)
    // alone.  Launch counts per tile phase (TilePhaseKind order).
    std::vector<int64_t> steps(n, 0), tile_launches(n, 0), ones(n, 0);
    std::vector<int32_t> launches(n, 0);
    std::vector<int64_t> cnt[kTilePhases];
    for (auto &c : cnt) c.assign(n, 0);
    for (int p = 0; p < n; ++p) {
        const nx_planet &q = planets[p];
        if (q.erosion_state) {
            int64_t cap = max_steps;
            if (q.step_cap > 0) cap = std::min<int64_t>(cap, q.step_cap);
            steps[p] = std::max<int64_t>(0, std::min<int64_t>(cap, q.erosion_iters - q.steps_done));
        }
        const int K = desc[p].has_plan ? desc[p].plan.steps : 0;
        const int64_t full = K ? steps[p] / K : 0;
        const int64_t rem = steps[p] - K * full;
        const int tail = (K == kMultiSteps && rem >= 2) ? (int)rem : 0;
        desc[p].full_launches = full;
        desc[p].tail_steps = tail;
        cnt[kPhaseTail3][p] = tail == 3;
        cnt[kPhaseTail2][p] = tail == 2;
        cnt[kPhaseQuad][p] = K == kMultiSteps ? full : 0;
        cnt[kPhasePair][p] = K == 2 ? full : 0;
        tile_launches[p] = full + (tail ? 1 : 0);
        ones[p] = rem - tail;
        launches[p] = (int32_t)(tile_launches[p] + ones[p]);
    }
    auto ordered = [&](const std::vector<int64_t> &c, std::vector<int32_t> &ord) {
        ord.resize(n);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return c[a] > c[b]; });
        int act = 0;
        while (act < n && c[ord[act]] > 0) ++act;
        return act;
    };
    // Each tile phase: planets ordered by launch count (the active set of
    // launch k is a prefix), their tile prefix, the bucket table of the
    // many-planet kernel (tile b in bucket b >> shift) and the kernel-
    // argument form of its first kArgPlanets planets.
    struct Phase {
        std::vector<int32_t> ord;
        std::vector<int64_t> tpre;
        std::vector<PairBucket> buckets;
        int active0 = 0, shift = 0;
        int64_t count0 = 0;
        TileLaunchArgs args;
    };
    std::vector<Phase> ph(kTilePhases);
    bool any = false;
    for (int f = 0; f < kTilePhases; ++f) {
        Phase &P = ph[f];
        P.active0 = ordered(cnt[f], P.ord);
        P.count0 = P.active0 ? cnt[f][P.ord[0]] : 0;
        any = any || P.active0 > 0;
        P.tpre.assign(n + 1, 0);
        for (int r = 0; r < n; ++r)
            P.tpre[r + 1] = P.tpre[r] + (r < P.active0 ? desc[P.ord[r]].plan.ntiles : 0);
        const int64_t nt = P.tpre[P.active0];
        if (nt > INT32_MAX) return set_error(NX_EINVAL, "%lld tiles in one batch exceed 2^31 - 1", (long long)nt);
        while (nt > 0 && ((nt - 1) >> P.shift) >= kPairBuckets) ++P.shift;
        P.buckets.assign(kPairBuckets, PairBucket{0, INT32_MAX});
        for (int j = 0, r = 0; j < kPairBuckets && ((int64_t)j << P.shift) < nt; ++j) {
            const int64_t first = (int64_t)j << P.shift;
            while (r + 1 < P.active0 && P.tpre[r + 1] <= first) ++r;
            P.buckets[j] = PairBucket{r, r + 1 < P.active0 ? (int32_t)P.tpre[r + 1] : INT32_MAX};
        }
        std::memset(&P.args, 0, sizeof(P.args));
        for (int r = 0; r < kArgPlanets; ++r) P.args.prefix[r] = INT32_MAX;
        for (int r = 0; r < std::min(P.active0, kArgPlanets); ++r) {
            const BatchPlanetDev &d = desc[P.ord[r]];
            P.args.prefix[r] = (int32_t)P.tpre[r];   // <= nt <= INT32_MAX, checked above
            P.args.pl[r] = TilePlanetArgs{d.plan, d.hw0,          d.hw1,  d.s0,        d.s1,
                                          d.meta, d.total_iters, d.rain, d.tail_steps};
        }
    }
    std::vector<int32_t> ordS;
    const int activeS0 = ordered(ones, ordS);
    if (!any && activeS0 == 0) return NX_OK;
    const int64_t KS = activeS0 ? ones[ordS[0]] : 0;
    std::vector<int64_t> bpre(n + 1, 0), sbase(n, 0);
    for (int r = 0; r < n; ++r) {
        bpre[r + 1] = bpre[r] + (r < activeS0 ? desc[ordS[r]].nblocks : 0);
        sbase[r] = tile_launches[ordS[r]];
    }
    char *w = (char *)ws;
    NX_HIP(hipSetDevice(ctx->device));
    std::vector<Part> parts = {{L.desc, desc.data(), sizeof(BatchPlanetDev) * n},
                               {L.one_prefix, bpre.data(), sizeof(int64_t) * (n + 1)},
                               {L.one_order, ordS.data(), sizeof(int32_t) * n},
                               {L.one_base, sbase.data(), sizeof(int64_t) * n},
                               {L.steps, steps.data(), sizeof(int64_t) * n},
                               {L.launches, launches.data(), sizeof(int32_t) * n}};
    for (int f = 0; f < kTilePhases; ++f) {
        parts.push_back({L.tile_prefix[f], ph[f].tpre.data(), sizeof(int64_t) * (n + 1)});
        parts.push_back({L.tile_order[f], ph[f].ord.data(), sizeof(int32_t) * n});
        parts.push_back({L.tile_bucket[f], ph[f].buckets.data(), sizeof(PairBucket) * kPairBuckets});
    }
    int rc = upload_parts(ctx, w, L, parts, stream);
    if (rc) return rc;
    const auto *D = (const BatchPlanetDev *)(w + L.desc);
    const auto *OS = (const int32_t *)(w + L.one_order);
    const auto *BP = (const int64_t *)(w + L.one_prefix);
    const auto *SB = (const int64_t *)(w + L.one_base);
    const auto *ST = (const int64_t *)(w + L.steps);
    const auto *LN = (const int32_t *)(w + L.launches);
    // Launch k of tile phase f: one planet, up to kArgPlanets, or more.
    auto tile_launch = [&](auto kernels, int f, int active, int64_t k, hipStream_t s) {
        const Phase &P = ph[f];
        const int64_t nt = P.tpre[active];
        const dim3 grid((unsigned)(((nt + 7) / 8) * 8));
        if (active == 1) {
            const BatchPlanetDev &d = desc[P.ord[0]];
            hipLaunchKernelGGL(kernels.one, grid, dim3(1024), 0, s, d.meta, d.hw0, d.hw1, d.s0,
                               d.s1, d.plan, d.tail_steps, k, d.total_iters, d.rain);
        } else if (active <= kArgPlanets) {
            // entries past the active planets start at or after nt
            hipLaunchKernelGGL(kernels.args, grid, dim3(1024), 0, s, P.args, (int32_t)nt, k);
        } else {
            hipLaunchKernelGGL(kernels.many, grid, dim3(1024), 0, s, D,
                               (const int32_t *)(w + L.tile_order[f]),
                               (const PairBucket *)(w + L.tile_bucket[f]), P.shift,
                               (const int64_t *)(w + L.tile_prefix[f]), active, nt, k);
        }
    };
    auto launch_all = [&](hipStream_t s) {
        // tails before the four-step launches of the same planet
        for (int f = 0; f < kTilePhases; ++f) {
            const Phase &P = ph[f];
            int active = P.active0;
            for (int64_t k = 0; k < P.count0; ++k) {
                while (active > 0 && cnt[f][P.ord[active - 1]] <= k) --active;
                switch (f) {
                case kPhaseTail3: tile_launch(TileKernels<4, 3>(), f, active, k, s); break;
                case kPhaseTail2: tile_launch(TileKernels<4, 2>(), f, active, k, s); break;
                case kPhaseQuad: tile_launch(TileKernels<4, 4>(), f, active, k, s); break;
                default: tile_launch(TileKernels<2, 2>(), f, active, k, s); break;
                }
            }
        }
        int active = activeS0;
        for (int64_t k = 0; k < KS; ++k) {
            while (active > 0 && ones[ordS[active - 1]] <= k) --active;
            const int64_t nb = bpre[active];
            hipLaunchKernelGGL(k_erode_batch, dim3((unsigned)(((nb + 7) / 8) * 8)), dim3(256), 0, s,
                               D, OS, BP, SB, active, nb, k);
        }
        hipLaunchKernelGGL(k_batch_advance, dim3((n + 255) / 256), dim3(256), 0, s, D, ST, LN, n);
    };
    if (!use_graph) {
        launch_all(stream);
        NX_HIP(hipGetLastError());
    } else {
        // Cache key: workspace, the descriptor contents and the schedule.
        uint64_t h = 1469598103934665603ull;
        h = fnv(desc.data(), sizeof(BatchPlanetDev) * n, h);
        for (const auto &c : cnt) h = fnv(c.data(), sizeof(int64_t) * n, h);
        h = fnv(ones.data(), sizeof(int64_t) * n, h);
        int64_t nl = KS;
        for (const Phase &P : ph) nl += P.count0;
        GraphKey key{ws, (int64_t)n, nullptr, nullptr, (int)nl, 0.0, true, h};
        hipGraphExec_t ex = nullptr;
        {
            std::lock_guard<std::mutex> lk(ctx->mu);
            auto itg = ctx->graphs.find(key);
            if (itg != ctx->graphs.end()) {
                ex = itg->second;
            } else {
                ctx->make_room_for_graph();
                hipGraph_t g = nullptr;
                NX_HIP(hipStreamBeginCapture(ctx->capture_stream, hipStreamCaptureModeThreadLocal));
                launch_all(ctx->capture_stream);
                const hipError_t le = hipGetLastError();
                const hipError_t ce = hipStreamEndCapture(ctx->capture_stream, &g);
                if (le != hipSuccess || ce != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    return set_error(NX_EHIP, "batch erosion capture failed: %s",
                                     hipGetErrorString(le != hipSuccess ? le : ce));
                }
                const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ie != hipSuccess) return set_error(NX_EHIP, "graph instantiate: %s", hipGetErrorString(ie));
                ctx->graphs.emplace(key, ex);
            }
        }
        NX_HIP(hipGraphLaunch(ex, stream));
    }
    for (int p = 0; p < n; ++p) planets[p].steps_done += steps[p];
    return NX_OK;
}

int batch_store(nx_ctx *ctx, const nx_planet *planets, int n, void *ws, hipStream_t stream)
{
    if (!ctx || !planets || n <= 0 || !ws) return set_error(NX_EINVAL, "batch needs planets and a workspace");
    const BatchLayout L = layout(n);
    std::vector<BatchPlanetDev> desc(n);
    std::vector<int64_t> rpre(n + 1, 0);
    for (int p = 0; p < n; ++p) {
        int rc = fill_desc(planets[p], desc[p]);
        if (rc) return rc;
        rpre[p + 1] = rpre[p] + (desc[p].V + kRsTile - 1) / kRsTile;
    }
    char *w = (char *)ws;
    NX_HIP(hipSetDevice(ctx->device));
    int rc = upload(ctx, w, L,
                    {{L.desc, desc.data(), sizeof(BatchPlanetDev) * n},
                     {L.rs_prefix, rpre.data(), sizeof(int64_t) * (n + 1)}},
                    stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_batch_store, dim3((unsigned)rpre[n]), dim3(256), 0, stream,
                       (const BatchPlanetDev *)(w + L.desc), (const int64_t *)(w + L.rs_prefix), n);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

}  // namespace nx
