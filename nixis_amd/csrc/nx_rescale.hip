// nx_rescale.hip -- the height post-processing after noise on the device:
// rescale in every form (util.py:110-175), the elevation mask
// (terrain.py:61-72), power_rescale (util.py:178-254) and the chain of
// nixis.py:330-364.
//
// Min/max reductions run per block and end in one atomic per block on an
// order-preserving integer encoding of the doubles; the maps then use the
// reference's FP64 operation order (IEEE division, no contraction), so
// rescale and the mask are bit-identical to the reference.  power_rescale's
// bounds follow the reference's sequential if/elif loop exactly, through an
// exclusive min-scan of the selected values (the running minimum each
// element was compared with).
#include <hip/hip_runtime.h>

#include <atomic>

#include "nx_ctx.h"
#include "nx_reduce_dev.h"

namespace nx {

// Grid-stride loop with four independent loads in flight per thread: all
// four values are loaded before the first is consumed (x and out may alias,
// so the compiler would not hoist a load above an earlier store itself).
template <class T, class Load, class Body>
__device__ __forceinline__ void stream4(int64_t n, Load load, Body body)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const T a = load(i), b = load(i + stride), c = load(i + 2 * stride),
                d = load(i + 3 * stride);
        body(i, a);
        body(i + stride, b);
        body(i + 2 * stride, c);
        body(i + 3 * stride, d);
    }
    for (; i < n; i += stride) body(i, load(i));
}

// The same with eight loads in flight per lane, for the read-only min/max.
template <typename T, typename Load, typename Body>
__device__ __forceinline__ void stream8(int64_t n, Load load, Body body)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 7 * stride < n; i += 8 * stride) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = load(i + u * stride);
#pragma unroll
        for (int u = 0; u < 8; ++u) body(i + u * stride, v[u]);
    }
    for (; i < n; i += stride) body(i, load(i));
}

// vec: the arrays are aligned for 16-byte accesses of value pairs; pairs
// (2j, 2j+1) stream with stream4 and block 0 thread 0 takes an odd last value.
__device__ __forceinline__ bool tail_thread() { return blockIdx.x == 0 && threadIdx.x == 0; }

__global__ void k_minmax_init(unsigned long long *keys)
{
    if (threadIdx.x == 0) {
        keys[0] = ~0ull;  // running min
        keys[1] = 0ull;   // running max
    }
}

__global__ __launch_bounds__(256) void k_minmax(const double *__restrict__ x, int64_t n,
                                                unsigned long long *__restrict__ keys, int vec)
{
    unsigned long long lo = ~0ull, hi = 0ull;
    auto upd = [&](double v) {
        const unsigned long long k = dkey(v);
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    };
    if (vec) {
        const double2 *x2 = (const double2 *)x;
        stream8<double2>(n >> 1, [&](int64_t j) { return x2[j]; },
                         [&](int64_t, double2 d) {
                             upd(d.x);
                             upd(d.y);
                         });
        if ((n & 1) && tail_thread()) upd(x[n - 1]);
    } else {
        stream8<double>(n, [&](int64_t i) { return x[i]; }, [&](int64_t, double v) { upd(v); });
    }
    block_minmax(lo, hi);
    if (threadIdx.x == 0) {
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
                                                     int *__restrict__ err, int vec)
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
    auto f = [&](double v) { return ((v - x_min) / x_range) * new_range + lower; };  // util.py:139
    if (vec) {
        const double2 *x2 = (const double2 *)x;
        double2 *o2 = (double2 *)out;
        stream4<double2>(n >> 1, [&](int64_t j) { return x2[j]; },
                         [&](int64_t j, double2 d) { o2[j] = make_double2(f(d.x), f(d.y)); });
        if ((n & 1) && tail_thread()) out[n - 1] = f(x[n - 1]);
    } else {
        stream4<double>(n, [&](int64_t i) { return x[i]; },
                        [&](int64_t i, double v) { out[i] = f(v); });
    }
}

static int vec16(const void *p) { return (uintptr_t)p % 16 == 0 ? 1 : 0; }
static int vec2(const void *p) { return (uintptr_t)p % 2 == 0 ? 1 : 0; }

// The read-only min/max runs one block per CU with eight loads in flight per
// lane: every block ends in two atomics on the same keys, and with eight
// blocks per CU those atomics at the end of the pass cost a large part of it
// (40 M values: 80 against 54 us alone, 89 against 79 us in the height chain).
static unsigned mm_grid(nx_ctx *ctx, int64_t n)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cus));
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
    const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cus * 8);
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, stream, keys);
    hipLaunchKernelGGL(k_minmax, dim3(mm_grid(ctx, n)), dim3(256), 0, stream, x, n, keys, vec16(x));
    hipLaunchKernelGGL(k_rescale_map, dim3(grid), dim3(256), 0, stream, x, n, keys, lower, upper,
                       out, minmax, ctx->d_err, vec16(x) & vec16(out));
    NX_HIP(hipGetLastError());
    return NX_OK;
}

// ---- rescale in every form (util.py:110-175) ------------------------------

struct RescaleArgs {
    double lower, upper, mid, u_min, u_max;
    int has_mid, mode, has_umin, has_umax;  // mode: 0 none, 1 'lower', 2 'upper'
};

// x and out may alias.  Values outside a one-sided mode's half are copied.
__global__ __launch_bounds__(256) void k_rescale_ex(const double *x, int64_t n,
                                                    const unsigned long long *__restrict__ keys,
                                                    RescaleArgs a, double *out,
                                                    double *__restrict__ minmax,
                                                    int *__restrict__ err, int vec)
{
    double x_min = dunkey(keys[0]), x_max = dunkey(keys[1]);
    if (blockIdx.x == 0 && threadIdx.x == 0 && minmax) {
        minmax[0] = x_min;
        minmax[1] = x_max;
    }
    if (a.has_umin && a.u_min < x_min) x_min = a.u_min;  // util.py:131-134
    if (a.has_umax && a.u_max > x_max) x_max = a.u_max;
    bool zero = false;
    auto f = [&](double v) {
        double r = v;
        if (a.mode == 0 && !a.has_mid) {
            const double xr = x_max - x_min;
            zero = zero | (xr == 0.0);
            r = ((v - x_min) / xr) * (a.upper - a.lower) + a.lower;
        } else if (a.mode == 0) {
            if (v <= a.mid) {
                const double xr = a.mid - x_min;
                zero = zero | (xr == 0.0);
                r = ((v - x_min) / xr) * (a.mid - a.lower) + a.lower;
            } else {
                const double xr = x_max - a.mid;
                zero = zero | (xr == 0.0);
                r = ((v - a.mid) / xr) * (a.upper - a.mid) + a.mid;
            }
        } else if (a.mode == 1 && v <= a.mid) {
            const double xr = a.mid - x_min;
            zero = zero | (xr == 0.0);
            r = ((v - x_min) / xr) * (a.mid - a.lower) + a.lower;
        } else if (a.mode == 2 && v >= a.mid) {
            const double xr = x_max - a.mid;
            zero = zero | (xr == 0.0);
            r = ((v - a.mid) / xr) * (a.upper - a.mid) + a.mid;
        }
        return r;
    };
    if (vec) {
        const double2 *x2 = (const double2 *)x;
        double2 *o2 = (double2 *)out;
        stream4<double2>(n >> 1, [&](int64_t j) { return x2[j]; },
                         [&](int64_t j, double2 d) { o2[j] = make_double2(f(d.x), f(d.y)); });
        if ((n & 1) && tail_thread()) out[n - 1] = f(x[n - 1]);
    } else {
        stream4<double>(n, [&](int64_t i) { return x[i]; },
                        [&](int64_t i, double v) { out[i] = f(v); });
    }
    if (__any(zero) && (threadIdx.x & 63) == 0)
        __hip_atomic_fetch_or(err, NX_ERR_DIV0_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

static unsigned red_grid(nx_ctx *ctx, int64_t n)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cus * 8));
}

static unsigned long long *key_slot(nx_ctx *ctx)
{
    return ctx->d_keys + 2 * (ctx->key_slot.fetch_add(1) % kKeySlots);
}

static int rescale_ex_launch(nx_ctx *ctx, const double *x, int64_t n, const RescaleArgs &a,
                             double *out, double *minmax, hipStream_t stream)
{
    unsigned long long *keys = key_slot(ctx);
    const unsigned grid = red_grid(ctx, n);
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, stream, keys);
    hipLaunchKernelGGL(k_minmax, dim3(mm_grid(ctx, n)), dim3(256), 0, stream, x, n, keys, vec16(x));
    hipLaunchKernelGGL(k_rescale_ex, dim3(grid), dim3(256), 0, stream, x, n, keys, a, out,
                       minmax, ctx->d_err, vec16(x) & vec16(out));
    NX_HIP(hipGetLastError());
    return NX_OK;
}

int rescale_ex(nx_ctx *ctx, const double *x, int64_t n, double lower, double upper,
               const nx_rescale_opts *o, double *out, double *minmax, hipStream_t stream)
{
    if (!ctx) return set_error(NX_EINVAL, "null context");
    if (n <= 0 || !x || !out) return set_error(NX_EINVAL, "rescale needs n > 0 and device arrays");
    RescaleArgs a{lower, upper, 0.0, 0.0, 0.0, 0, 0, 0, 0};
    if (o) {
        if (o->mode < 0 || o->mode > 2) return set_error(NX_EINVAL, "rescale mode %d outside 0..2", o->mode);
        a.mid = o->mid;
        a.u_min = o->u_min;
        a.u_max = o->u_max;
        a.has_mid = o->has_mid != 0;
        a.mode = o->mode;
        a.has_umin = o->has_umin != 0;
        a.has_umax = o->has_umax != 0;
    }
    if (a.mode != 0 && !a.has_mid)
        return set_error(NX_EINVAL, "rescale: a mode needs mid (the reference returns x unchanged)");
    NX_HIP(hipSetDevice(ctx->device));
    return rescale_ex_launch(ctx, x, n, a, out, minmax, stream);
}

// ---- terrain.py:61-72 ---------------------------------------------------------

__global__ __launch_bounds__(256) void k_mask(const double *__restrict__ h, int64_t n,
                                              const double *__restrict__ level_dev, double level,
                                              uint8_t *__restrict__ mask, int vec)
{
    const double lv = level_dev ? *level_dev : level;
    if (vec) {
        const double2 *h2 = (const double2 *)h;
        uint16_t *m2 = (uint16_t *)mask;
        stream4<double2>(n >> 1, [&](int64_t j) { return h2[j]; }, [&](int64_t j, double2 d) {
            m2[j] = (uint16_t)((d.x <= lv ? 1 : 0) | (d.y <= lv ? 0x100 : 0));
        });
        if ((n & 1) && tail_thread()) mask[n - 1] = h[n - 1] <= lv ? 1 : 0;
    } else {
        stream4<double>(n, [&](int64_t i) { return h[i]; },
                        [&](int64_t i, double v) { mask[i] = v <= lv ? 1 : 0; });
    }
}

int make_mask(nx_ctx *ctx, const double *h, int64_t n, double level, uint8_t *mask,
              hipStream_t stream)
{
    if (!ctx || !h || !mask || n <= 0) return set_error(NX_EINVAL, "mask needs n > 0 and device arrays");
    NX_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_mask, dim3(red_grid(ctx, n)), dim3(256), 0, stream, h, n,
                       (const double *)nullptr, level, mask, vec16(h) & vec2(mask));
    NX_HIP(hipGetLastError());
    return NX_OK;
}

// ---- power_rescale (util.py:178-254) ----------------------------------------
//
// The reference finds the bounds of the selected values with one sequential
// loop, starting from lower = max(x), upper = min(x):
//   if sel and x < lower: lower = x   elif sel and x > upper: upper = x
// so a value that lowers `lower` (a running minimum) is never compared with
// `upper`.  Let s_0 = max(x) and s_1, s_2, ... be the selected values in
// order (a NaN changes neither bound and is left out).  Up to the first j
// with s_j >= s_{j-1} every s is a new running minimum; s_j is not, and every
// later running minimum is below s_j, so the largest value from s_j on is
// never a running minimum.  Hence
//   lower = min(max(x), min s)
//   upper = max(min(x), max(s_j, s_{j+1}, ...))   (min(x) when there is no j).
// Both follow from an associative summary of a run of selected values (first,
// last, min, max, whether the run holds such a break and the maximum from its
// first break on), reduced in order per chunk and then across chunks: one
// read of x and the mask instead of a scan.

struct PSum {
    double first, last, mn, mx, suf;
    int ne, brk;  // any value; a break inside the run
};

// Device scalars at the head of the power_rescale / height-chain workspace.
struct PrHead {
    unsigned long long keys0[2], keys1[2], keys2[2];  // min/max reduction keys
    double xmin, xmax;  // min/max of the array the bounds start from
    double level;       // ocean level (height chain)
    double pad;
    PSum s[2];          // whole-array summaries: [0] mask unset, [1] mask set
};
static_assert(sizeof(PrHead) <= 256, "PrHead must fit the 256-byte workspace head");

constexpr int kPrThreads = 256;
constexpr int kPrPer = 8;                       // consecutive values per lane
constexpr int64_t kPrUnit = (int64_t)64 * kPrPer;  // values per wave chunk

__device__ __forceinline__ PSum psum_empty()
{
    PSum s;
    s.first = s.last = s.mn = s.mx = s.suf = 0.0;
    s.ne = s.brk = 0;
    return s;
}

// Append one (non-NaN) value to a run.
__device__ __forceinline__ void psum_push(PSum &s, double v)
{
    if (!s.ne) {
        s.first = s.last = s.mn = s.mx = v;
        s.ne = 1;
        return;
    }
    if (s.brk)
        s.suf = fmax(s.suf, v);
    else if (!(v < s.last)) {
        s.brk = 1;
        s.suf = v;
    }
    s.last = v;
    s.mn = fmin(s.mn, v);
    s.mx = fmax(s.mx, v);
}

// Run a followed by run b.
__device__ __forceinline__ PSum psum_join(const PSum &a, const PSum &b)
{
    if (!a.ne) return b;
    if (!b.ne) return a;
    PSum r;
    r.ne = 1;
    r.first = a.first;
    r.last = b.last;
    r.mn = fmin(a.mn, b.mn);
    r.mx = fmax(a.mx, b.mx);
    if (a.brk) {
        r.brk = 1;
        r.suf = fmax(a.suf, b.mx);
    } else if (!(b.first < a.last)) {
        r.brk = 1;
        r.suf = b.mx;
    } else {
        r.brk = b.brk;
        r.suf = b.suf;
    }
    return r;
}

__device__ __forceinline__ PSum psum_shfl_down(const PSum &s, int off)
{
    PSum r;
    r.first = __shfl_down(s.first, off, 64);
    r.last = __shfl_down(s.last, off, 64);
    r.mn = __shfl_down(s.mn, off, 64);
    r.mx = __shfl_down(s.mx, off, 64);
    r.suf = __shfl_down(s.suf, off, 64);
    r.ne = __shfl_down(s.ne, off, 64);
    r.brk = __shfl_down(s.brk, off, 64);
    return r;
}

// In-order reduction over one wave; the result is valid in lane 0.
__device__ __forceinline__ PSum psum_wave(PSum s)
{
    const int lane = threadIdx.x & 63;
    for (int off = 1; off < 64; off <<= 1) {
        const PSum o = psum_shfl_down(s, off);
        if ((lane & (2 * off - 1)) == 0) s = psum_join(s, o);
    }
    return s;
}

// In-order reduction over the block (lane order within waves, then wave
// order); the result is valid in thread 0.  lds holds one entry per wave.
template <int NW>
__device__ PSum psum_block(PSum s, PSum *lds)
{
    s = psum_wave(s);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < NW; ++w) s = psum_join(s, lds[w]);
    __syncthreads();  // lds is reused by the next call
    return s;
}

// Bounds of the selected values from the whole array's summary.
__device__ __forceinline__ void psum_bounds(const PSum &s, double x_min, double x_max, double &lo,
                                            double &up)
{
    lo = x_max;
    up = x_min;
    if (!s.ne) return;
    if (s.mn < lo) lo = s.mn;
    bool has = true;
    double cand = 0.0;
    if (!(s.first < x_max))
        cand = s.mx;  // s_1 itself is no running minimum
    else if (s.brk)
        cand = s.suf;
    else
        has = false;
    if (has && cand > up) up = cand;
}

struct ChainArgs {
    double lower, upper, percent;  // first rescale and ocean percent
    double *level_out;             // device f64[1]
};

// Summaries of the values whose mask is unset (sum[2b]) and set (sum[2b + 1])
// over block b's values.  Wave w of block b takes the contiguous run of
// wave chunks [(4b + w) * cpw, (4b + w + 1) * cpw), kPrPer consecutive values
// per lane, and keeps its running summary in lane 0 (no barrier in the
// loop); the four waves join in order at the end.  CHAIN: x holds the noise
// heights; each value is rewritten as the standard rescale (util.py:139)
// with min/max from head->keys0, the rescaled min/max and the ocean level
// are derived from the same keys (the map is monotone, so its extremes are
// the images of x's), and mask = t <= level (terrain.py:61-72) is written.
// vec: x is 16-byte and mask 8-byte aligned (vector accesses).
template <bool CHAIN>
__global__ __launch_bounds__(kPrThreads) void k_pr_summary(double *x, uint8_t *mask, int64_t n,
                                                           int64_t nunits, int64_t cpw, int vec,
                                                           PSum *__restrict__ sum,
                                                           PrHead *__restrict__ head, ChainArgs ca,
                                                           int *__restrict__ err)
{
    __shared__ PSum lds[2][kPrThreads / 64];
    double x_min = 0.0, xr = 1.0, nr = 0.0, level = 0.0;
    if (CHAIN) {
        x_min = dunkey(head->keys0[0]);
        const double x_max = dunkey(head->keys0[1]);
        xr = x_max - x_min;
        nr = ca.upper - ca.lower;
        const double fa = ((x_min - x_min) / xr) * nr + ca.lower;
        const double fb = ((x_max - x_min) / xr) * nr + ca.lower;
        const double tmin = fa < fb ? fa : fb, tmax = fa < fb ? fb : fa;
        level = tmin + ((tmax - tmin) * ca.percent / 100.0);  // util.py:556-566
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            head->xmin = tmin;
            head->xmax = tmax;
            head->level = level;
            *ca.level_out = level;
            if (xr == 0.0)
                __hip_atomic_fetch_or(err, NX_ERR_DIV0_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    PSum r0 = psum_empty(), r1 = psum_empty();  // the wave's running summaries (lane 0)
    const int64_t u_begin = ((int64_t)blockIdx.x * (kPrThreads / 64) + wave) * cpw;
    const int64_t u_end = u_begin + cpw < nunits ? u_begin + cpw : nunits;
    // Values of wave chunk u (and their mask bits unless CHAIN); lanes past
    // the end get cnt = 0.
    auto load = [&](int64_t u, double *v, bool *m, int &cnt) {
        const int64_t base = u * kPrUnit + (int64_t)lane * kPrPer;
        const int64_t left = n - base;
        cnt = left <= 0 ? 0 : (left >= kPrPer ? kPrPer : (int)left);
        if (cnt == kPrPer && vec) {
            const double2 *x2 = (const double2 *)(x + base);
#pragma unroll
            for (int k = 0; k < kPrPer / 2; ++k) {
                const double2 t = x2[k];
                v[2 * k] = t.x;
                v[2 * k + 1] = t.y;
            }
            if (!CHAIN) {
                const uint64_t w = *(const uint64_t *)(mask + base);
#pragma unroll
                for (int k = 0; k < kPrPer; ++k) m[k] = ((w >> (8 * k)) & 0xff) != 0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < kPrPer; ++k) {
                v[k] = k < cnt ? x[base + k] : 0.0;
                if (!CHAIN) m[k] = k < cnt ? mask[base + k] != 0 : false;
            }
        }
    };
    double vn[kPrPer];
    bool mn[kPrPer] = {};
    int cn = 0;
    if (u_begin < u_end) load(u_begin, vn, mn, cn);
    for (int64_t u = u_begin; u < u_end; ++u) {
        const int64_t base = u * kPrUnit + (int64_t)lane * kPrPer;
        double v[kPrPer];
        bool m[kPrPer];
        const int cnt = cn;
#pragma unroll
        for (int k = 0; k < kPrPer; ++k) {
            v[k] = vn[k];
            m[k] = mn[k];
        }
        if (u + 1 < u_end) load(u + 1, vn, mn, cn);  // in flight while u is processed
        if (CHAIN) {
#pragma unroll
            for (int k = 0; k < kPrPer; ++k) {
                v[k] = ((v[k] - x_min) / xr) * nr + ca.lower;
                m[k] = v[k] <= level;
            }
            if (cnt == kPrPer && vec) {
                double2 *x2 = (double2 *)(x + base);
                uint64_t w = 0;
#pragma unroll
                for (int k = 0; k < kPrPer / 2; ++k) x2[k] = make_double2(v[2 * k], v[2 * k + 1]);
#pragma unroll
                for (int k = 0; k < kPrPer; ++k) w |= (uint64_t)(m[k] ? 1 : 0) << (8 * k);
                *(uint64_t *)(mask + base) = w;
            } else {
#pragma unroll
                for (int k = 0; k < kPrPer; ++k)
                    if (k < cnt) {
                        x[base + k] = v[k];
                        mask[base + k] = m[k] ? 1 : 0;
                    }
            }
        }
        PSum s0 = psum_empty(), s1 = psum_empty();
#pragma unroll
        for (int k = 0; k < kPrPer; ++k)
            if (k < cnt && v[k] == v[k]) {
                if (m[k])
                    psum_push(s1, v[k]);
                else
                    psum_push(s0, v[k]);
            }
        s0 = psum_wave(s0);
        s1 = psum_wave(s1);
        if (lane == 0) {
            r0 = psum_join(r0, s0);
            r1 = psum_join(r1, s1);
        }
    }
    if (lane == 0) {
        lds[0][wave] = r0;
        lds[1][wave] = r1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kPrThreads / 64; ++w) {
            r0 = psum_join(r0, lds[0][w]);
            r1 = psum_join(r1, lds[1][w]);
        }
        sum[2 * blockIdx.x] = r0;
        sum[2 * blockIdx.x + 1] = r1;
    }
}

// One block: in-order reduction of the per-block summaries into head->s;
// with keys, head->xmin/xmax = (min, max) of x from them.
__global__ __launch_bounds__(1024) void k_pr_finish(const PSum *__restrict__ sum, int64_t nsum,
                                                    PrHead *__restrict__ head,
                                                    const unsigned long long *__restrict__ keys)
{
    __shared__ PSum lds[2][16];
    const int64_t per = (nsum + blockDim.x - 1) / blockDim.x;
    const int64_t c0 = (int64_t)threadIdx.x * per;
    const int64_t c1 = c0 + per < nsum ? c0 + per : nsum;
    PSum s0 = psum_empty(), s1 = psum_empty();
    for (int64_t c = c0; c < c1; ++c) {
        s0 = psum_join(s0, sum[2 * c]);
        s1 = psum_join(s1, sum[2 * c + 1]);
    }
    s0 = psum_block<16>(s0, lds[0]);
    s1 = psum_block<16>(s1, lds[1]);
    if (threadIdx.x == 0) {
        head->s[0] = s0;
        head->s[1] = s1;
        if (keys) {
            head->xmin = dunkey(keys[0]);
            head->xmax = dunkey(keys[1]);
        }
    }
}

__device__ __forceinline__ bool pr_sel(int mode, uint8_t m) { return mode == 1 ? m != 0 : (mode == 0 && m == 0); }

// util.py:224-248 for the selected values; the others are copied.  The
// bounds start from (min, max) of the input: from keys when given, else from
// head->xmin/xmax.  shift: subtract head->level from every value afterwards
// (nixis.py:362).  okeys (nullable): min/max keys of the values written.
// x and out may alias.  vec: x and out 16-byte, mask 2-byte aligned.
__global__ __launch_bounds__(256) void k_pr_map(const double *x, double *out,
                                                const uint8_t *__restrict__ mask, int64_t n,
                                                int mode, double power,
                                                const PrHead *__restrict__ head,
                                                const unsigned long long *__restrict__ keys,
                                                int shift, unsigned long long *__restrict__ okeys,
                                                int *__restrict__ err, int vec)
{
    const double x_min = keys ? dunkey(keys[0]) : head->xmin;
    const double x_max = keys ? dunkey(keys[1]) : head->xmax;
    double mask_lower = x_max, mask_upper = x_min;
    if (mode == 0 || mode == 1) psum_bounds(head->s[mode], x_min, x_max, mask_lower, mask_upper);
    const double level = shift ? head->level : 0.0;
    const double temp_upper = 1.0, temp_lower = 0.0, temp_range = temp_upper - temp_lower;
    const double mask_range = mask_upper - mask_lower;
    bool zero = false;
    unsigned long long lo = ~0ull, hi = 0ull;
    auto f = [&](double r, uint8_t m) {
        if (pr_sel(mode, m)) {
            zero = zero | (mask_range == 0.0);
            double t = ((r - mask_lower) / mask_range) * temp_range + temp_lower;
            // x ** 2 and x ** 0.5 as the correctly rounded t * t and sqrt(t),
            // the values an accurate pow returns (the height chain's powers).
            if (power == 2.0)
                t = t * t;
            else if (power == 0.5)
                t = t == 0.0 ? 0.0 : sqrt(t);  // pow(-0, 0.5) is +0
            else
                t = pow(t, power);
            r = ((t - temp_lower) / temp_range) * mask_range + mask_lower;
        }
        if (shift) r = r - level;
        if (okeys) {
            const unsigned long long k = dkey(r);
            lo = k < lo ? k : lo;
            hi = k > hi ? k : hi;
        }
        return r;
    };
    if (vec) {
        struct VM2 {
            double2 v;
            uint16_t m;
        };
        const double2 *x2 = (const double2 *)x;
        const uint16_t *m2 = (const uint16_t *)mask;
        double2 *o2 = (double2 *)out;
        stream4<VM2>(n >> 1, [&](int64_t j) { return VM2{x2[j], m2[j]}; }, [&](int64_t j, VM2 e) {
            const double r0 = f(e.v.x, (uint8_t)(e.m & 0xff));
            const double r1 = f(e.v.y, (uint8_t)(e.m >> 8));
            o2[j] = make_double2(r0, r1);
        });
        if ((n & 1) && tail_thread()) out[n - 1] = f(x[n - 1], mask[n - 1]);
    } else {
        struct VM {
            double v;
            uint8_t m;
        };
        stream4<VM>(n, [&](int64_t i) { return VM{x[i], mask[i]}; },
                    [&](int64_t i, VM e) { out[i] = f(e.v, e.m); });
    }
    if (__any(zero) && (threadIdx.x & 63) == 0)
        __hip_atomic_fetch_or(err, NX_ERR_DIV0_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (okeys) {
        block_minmax(lo, hi);
        if (threadIdx.x == 0) {
            __hip_atomic_fetch_min(&okeys[0], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_max(&okeys[1], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

// Summary blocks: at most kPrMaxBlocks; each wave takes a contiguous run of
// wave chunks.
constexpr int64_t kPrMaxBlocks = 2048;

struct PrGrid {
    int64_t nunits, cpw, blocks;
};

static PrGrid pr_grid(nx_ctx *ctx, int64_t n)
{
    constexpr int64_t wpb = kPrThreads / 64;
    PrGrid g;
    g.nunits = (n + kPrUnit - 1) / kPrUnit;
    const int64_t want = wpb * std::max<int64_t>(1, std::min<int64_t>(kPrMaxBlocks, (int64_t)ctx->num_cus * 8));
    g.cpw = (g.nunits + want - 1) / want;
    const int64_t waves = (g.nunits + g.cpw - 1) / g.cpw;
    g.blocks = (waves + wpb - 1) / wpb;
    return g;
}

size_t power_rescale_workspace_bytes(int64_t n)
{
    return 256 + (n > 0 ? a256(sizeof(PSum) * 2 * (size_t)kPrMaxBlocks) : 0);
}

static int pr_vec(const void *x, const void *mask)
{
    return ((uintptr_t)x % 16 == 0 && (uintptr_t)mask % 8 == 0) ? 1 : 0;
}

int power_rescale(nx_ctx *ctx, const double *x, const uint8_t *mask, int64_t n, int mode,
                  double power, double *out, void *ws, hipStream_t stream)
{
    if (!ctx || !x || !mask || !out || !ws || n <= 0)
        return set_error(NX_EINVAL, "power_rescale needs device arrays, a workspace and n > 0");
    if (((uintptr_t)ws & 255) != 0) return set_error(NX_EINVAL, "workspace must be 256-byte aligned");
    NX_HIP(hipSetDevice(ctx->device));
    auto *head = (PrHead *)ws;
    auto *sum = (PSum *)((char *)ws + 256);
    const PrGrid g = pr_grid(ctx, n);
    const unsigned grid = red_grid(ctx, n);
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, stream, head->keys0);
    hipLaunchKernelGGL(k_minmax, dim3(mm_grid(ctx, n)), dim3(256), 0, stream, x, n, head->keys0,
                       vec16(x));
    hipLaunchKernelGGL(k_pr_summary<false>, dim3((unsigned)g.blocks), dim3(kPrThreads), 0, stream,
                       const_cast<double *>(x), const_cast<uint8_t *>(mask), n, g.nunits, g.cpw,
                       pr_vec(x, mask), sum, head, ChainArgs{0.0, 0.0, 0.0, nullptr}, ctx->d_err);
    hipLaunchKernelGGL(k_pr_finish, dim3(1), dim3(1024), 0, stream, sum, g.blocks, head,
                       (const unsigned long long *)head->keys0);
    hipLaunchKernelGGL(k_pr_map, dim3(grid), dim3(256), 0, stream, x, out, mask, n, mode, power,
                       head, (const unsigned long long *)nullptr, 0,
                       (unsigned long long *)nullptr, ctx->d_err, vec16(x) & vec16(out) & vec2(mask));
    NX_HIP(hipGetLastError());
    return NX_OK;
}

// ---- nixis.py:330-364: the chain after noise ------------------------------------
//
// Five passes over the heights: min/max; rescale + mask + power_rescale
// summaries of both halves (the land half's values are not changed by the
// ocean half's power_rescale, so one pass summarises both); the ocean
// power_rescale with the min/max of its result; the land power_rescale with
// the shift by the ocean level and the min/max of its result; the final
// split rescale around 0.

size_t height_chain_workspace_bytes(int64_t n) { return power_rescale_workspace_bytes(n); }

// The chain after head->keys0 holds the min/max keys of h.
static int chain_after_keys(nx_ctx *ctx, double *h, int64_t n, double min_alt, double max_alt,
                            double ocean_percent, uint8_t *mask, double *level, void *ws,
                            hipStream_t stream)
{
    if (!(0.0 < ocean_percent && ocean_percent < 100.0)) ocean_percent = 50.0;  // util.py:561-564
    auto *head = (PrHead *)ws;
    auto *sum = (PSum *)((char *)ws + 256);
    const PrGrid g = pr_grid(ctx, n);
    const unsigned grid = red_grid(ctx, n);
    // nixis.py:332-346 in one pass
    hipLaunchKernelGGL(k_pr_summary<true>, dim3((unsigned)g.blocks), dim3(kPrThreads), 0, stream,
                       h, mask, n, g.nunits, g.cpw, pr_vec(h, mask), sum, head,
                       ChainArgs{min_alt, max_alt, ocean_percent, level}, ctx->d_err);
    hipLaunchKernelGGL(k_pr_finish, dim3(1), dim3(1024), 0, stream, sum, g.blocks, head,
                       (const unsigned long long *)nullptr);
    // nixis.py:352: the ocean floor, power 0.5
    const int vmap = vec16(h) & vec2(mask);
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, stream, head->keys1);
    hipLaunchKernelGGL(k_pr_map, dim3(grid), dim3(256), 0, stream, h, h, mask, n, 1, 0.5, head,
                       (const unsigned long long *)nullptr, 0, head->keys1, ctx->d_err, vmap);
    // nixis.py:355 and 362: the land, power 2.0, then the shift
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, stream, head->keys2);
    hipLaunchKernelGGL(k_pr_map, dim3(grid), dim3(256), 0, stream, h, h, mask, n, 0, 2.0, head,
                       (const unsigned long long *)head->keys1, 1, head->keys2, ctx->d_err, vmap);
    // nixis.py:364
    RescaleArgs mid_args{min_alt, max_alt, 0.0, 0.0, 0.0, 1, 0, 0, 0};
    hipLaunchKernelGGL(k_rescale_ex, dim3(grid), dim3(256), 0, stream, h, n,
                       (const unsigned long long *)head->keys2, mid_args, h, (double *)nullptr,
                       ctx->d_err, vec16(h));
    NX_HIP(hipGetLastError());
    return NX_OK;
}

static int chain_check(nx_ctx *ctx, const void *h, int64_t n, const void *mask, const void *level,
                       const void *ws)
{
    if (!ctx || !h || !mask || !level || !ws || n <= 0)
        return set_error(NX_EINVAL, "height chain needs device arrays, a workspace and n > 0");
    if (((uintptr_t)ws & 255) != 0) return set_error(NX_EINVAL, "workspace must be 256-byte aligned");
    return NX_OK;
}

int height_chain(nx_ctx *ctx, double *h, int64_t n, double min_alt, double max_alt,
                 double ocean_percent, uint8_t *mask, double *level, void *ws, hipStream_t stream)
{
    int rc = chain_check(ctx, h, n, mask, level, ws);
    if (rc) return rc;
    NX_HIP(hipSetDevice(ctx->device));
    auto *head = (PrHead *)ws;
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, stream, head->keys0);
    hipLaunchKernelGGL(k_minmax, dim3(mm_grid(ctx, n)), dim3(256), 0, stream, h, n, head->keys0,
                       vec16(h));
    return chain_after_keys(ctx, h, n, min_alt, max_alt, ocean_percent, mask, level, ws, stream);
}

// terrain.py:32-59 into h (starting from zero, as nixis.py:330 passes no
// elevations) with the min/max keys reduced by the noise kernel, then the chain.
int terrain_heights(nx_ctx *ctx, const double *verts, int64_t V, const int32_t *perm,
                    const int32_t *pgi, int n_oct, const double *rough, const double *strength,
                    double radius, double min_alt, double max_alt, double ocean_percent,
                    double *h, uint8_t *mask, double *level, void *ws, hipStream_t stream)
{
    int rc = chain_check(ctx, h, V, mask, level, ws);
    if (rc) return rc;
    NX_HIP(hipSetDevice(ctx->device));
    auto *head = (PrHead *)ws;
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, stream, head->keys0);
    rc = sample_octaves_keys(ctx, verts, V, perm, pgi, n_oct, rough, strength, radius, h, 0,
                             nullptr, nullptr, head->keys0, stream);
    if (rc) return rc;
    return chain_after_keys(ctx, h, V, min_alt, max_alt, ocean_percent, mask, level, ws, stream);
}

// ---- terrain_heights without a caller workspace ---------------------------
//
// The chain's workspace (the 256-byte head and at most kPrMaxBlocks pairs of
// block summaries) as a file-scope device array: one copy per GPU, held
// until the process ends.  Its device address is taken once per context,
// with the context's device current (chain_scratch_init, called when the
// context is created), and passed to the kernels exactly as terrain_heights
// passes a caller's workspace.  Every reduction key in the head is reset by a
// fill kernel at the start of its pass and every summary is written before
// it is read, so nothing carries over from the previous call; calls that
// overlap on one GPU share the array and race.
constexpr size_t kChainScratchBytes = 256 + ((sizeof(PSum) * 2 * (size_t)kPrMaxBlocks + 255) & ~(size_t)255);
static_assert(kChainScratchBytes == 196864, "include/nixis.h states the scratch size");
static __device__ __align__(256) unsigned char scratch[kChainScratchBytes];

int chain_scratch_init(nx_ctx *ctx)
{
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(NX_EHIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    ctx->chain_scratch = p;
    return NX_OK;
}

int terrain_heights_noalloc(nx_ctx *ctx, const double *verts, int64_t V, const int32_t *perm,
                            const int32_t *pgi, int n_oct, const double *rough,
                            const double *strength, double radius, double min_alt,
                            double max_alt, double ocean_percent, double *h, uint8_t *mask,
                            double *level, hipStream_t stream)
{
    if (!ctx) return set_error(NX_EINVAL, "null context");
    if (!ctx->chain_scratch) return set_error(NX_EINVAL, "context has no static scratch address");
    const size_t need = height_chain_workspace_bytes(V);
    if (need > kChainScratchBytes)
        return set_error(NX_ERANGE, "V=%lld needs %zu bytes of chain scratch; the static scratch holds %zu",
                         (long long)V, need, kChainScratchBytes);
    return terrain_heights(ctx, verts, V, perm, pgi, n_oct, rough, strength, radius, min_alt,
                           max_alt, ocean_percent, h, mask, level, ctx->chain_scratch, stream);
}

}  // namespace nx
