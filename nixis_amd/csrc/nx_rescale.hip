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

// vec: the arrays are aligned for 16-byte accesses of value pairs; pairs
// (2j, 2j+1) stream with stream4 and block 0 thread 0 takes an odd last value.
__device__ __forceinline__ bool tail_thread() { return blockIdx.x == 0 && threadIdx.x == 0; }

constexpr int kMinmaxThreads = 1024;

__global__ void k_minmax_init(unsigned long long *keys)
{
    if (threadIdx.x == 0) {
        keys[0] = ~0ull;  // running min
        keys[1] = 0ull;   // running max
    }
}

// Min/max keys of x.  Each wave takes a contiguous run of 512-value chunks
// (four slices of 128, lane l holding values 2l and 2l + 1 of each, so every
// load instruction covers one contiguous range) with the next chunk's loads
// in flight while it reduces the current one.  vec: x is 16-byte aligned.
__global__ __launch_bounds__(kMinmaxThreads) void k_minmax(const double *__restrict__ x, int64_t n,
                                                unsigned long long *__restrict__ keys, int vec)
{
    unsigned long long lo = ~0ull, hi = 0ull;
    auto upd = [&](double v) {
        const unsigned long long k = dkey(v);
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    };
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t wid = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t nunits = (n + 511) / 512;
    const int64_t cpw = (nunits + waves - 1) / waves;
    const int64_t u0 = wid * cpw, u1 = u0 + cpw < nunits ? u0 + cpw : nunits;
    // value j of the lane's share of chunk u sits at u * 512 + 2 lane + off(j)
    auto off = [](int j) { return (int64_t)(j >> 1) * 128 + (j & 1); };
    auto load = [&](int64_t u, double *v, int64_t &left) {
        const int64_t p0 = u * 512 + 2 * lane;
        left = n - p0;
        if (vec && off(7) < left) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double2 t = *(const double2 *)(x + p0 + k * 128);
                v[2 * k] = t.x;
                v[2 * k + 1] = t.y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = off(j) < left ? x[p0 + off(j)] : 0.0;
        }
    };
    double vn[8];
    int64_t leftn = 0;
    if (u0 < u1) load(u0, vn, leftn);
    for (int64_t u = u0; u < u1; ++u) {
        double v[8];
        const int64_t left = leftn;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = vn[j];
        if (u + 1 < u1) load(u + 1, vn, leftn);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (off(j) < left) upd(v[j]);
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
// One block of kMinmaxThreads per CU: sixteen waves streaming, two atomics
// per CU.
static unsigned mm_grid(nx_ctx *ctx, int64_t n)
{
    return (unsigned)std::max<int64_t>(
        1, std::min<int64_t>((n + kMinmaxThreads - 1) / kMinmaxThreads, (int64_t)ctx->num_cus));
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
    hipLaunchKernelGGL(k_minmax, dim3(mm_grid(ctx, n)), dim3(kMinmaxThreads), 0, stream, x, n, keys, vec16(x));
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
                                                    int *__restrict__ err, int vec,
                                                    const int32_t *__restrict__ skip)
{
    if (skip && *skip) return;
    double x_min = dunkey(keys[0]), x_max = dunkey(keys[1]);
    // numba's np.min and np.max both return the array's first NaN, so one
    // NaN makes both extremes NaN (the keys put a NaN at one end only).
    if (isnan(x_min) || isnan(x_max)) x_min = x_max = isnan(x_min) ? x_min : x_max;
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
    hipLaunchKernelGGL(k_minmax, dim3(mm_grid(ctx, n)), dim3(kMinmaxThreads), 0, stream, x, n, keys, vec16(x));
    hipLaunchKernelGGL(k_rescale_ex, dim3(grid), dim3(256), 0, stream, x, n, keys, a, out,
                       minmax, ctx->d_err, vec16(x) & vec16(out), (const int32_t *)nullptr);
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
    // Height chain in two passes (derived by k_pr_finish, read by
    // k_chain_final): the ocean and land power_rescale bounds (lower, range)
    // and the min/max of the shifted heights, and whether they were derived
    // (else the five-pass kernels run).
    double lo1, r1, lo0, r0, h4min, h4max;
    int32_t fast, pad2;
};
static_assert(sizeof(PrHead) <= 256, "PrHead must fit the 256-byte workspace head");
// The flag's byte offset in the workspace is part of the tested behaviour
// (tests read it to see which form ran).
static_assert(offsetof(PrHead, fast) == 224, "PrHead::fast moved");

constexpr int kPrThreads = 512;
constexpr int kFinalThreads = 256;  // the block size red_grid sizes grids for
constexpr int kPrPer = 8;                          // values per lane per wave chunk
constexpr int kPrSlices = kPrPer / 2;              // slices of 128 values per chunk
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

// util.py:224-248 on one selected value r, with the selected values' lower
// bound and range.  x ** 2 and x ** 0.5 as the correctly rounded t * t and
// sqrt(t), the values an accurate pow returns (the height chain's powers).
__device__ __forceinline__ double pr_power_map(double r, double mask_lower, double mask_range,
                                               double power)
{
    const double temp_upper = 1.0, temp_lower = 0.0, temp_range = temp_upper - temp_lower;
    double t = ((r - mask_lower) / mask_range) * temp_range + temp_lower;
    if (power == 2.0)
        t = t * t;
    else if (power == 0.5)
        t = t == 0.0 ? 0.0 : sqrt(t);  // pow(-0, 0.5) is +0
    else
        t = pow(t, power);
    return ((t - temp_lower) / temp_range) * mask_range + mask_lower;
}

// nixis.py:364 on one value: rescale(h, lower, upper, mid=0) (util.py:141-160
// with mid), x_min / x_max the array's; zero reports a zero range used.
__device__ __forceinline__ double split_rescale(double v, double x_min, double x_max,
                                                double lower, double upper, bool &zero)
{
    const double mid = 0.0;
    if (v <= mid) {
        const double xr = mid - x_min;
        zero = zero | (xr == 0.0);
        return ((v - x_min) / xr) * (mid - lower) + lower;
    }
    const double xr = x_max - mid;
    zero = zero | (xr == 0.0);
    return ((v - mid) / xr) * (upper - mid) + mid;
}

// The height chain's constants after its summary pass (nixis.py:348-364).
struct ChainConsts {
    double lo1, r1, lo0, r0, h4min, h4max;  // ocean / land bounds; shifted min/max
    int fast;                               // all derived (else the five-pass form)
};

// The constants from the whole-array summaries so (mask set, the ocean) and
// sl (mask unset, the land) and the rescaled heights' min/max and level, in
// one thread: the ocean bounds from the ocean summary and the min/max of the
// rescaled heights; the min/max after the ocean map as the images of the
// ocean values' extremes beside the land values' own; the land bounds from
// those; the min/max after the land map and the shift likewise.  Each map
// is monotone non-decreasing on its selected values when its range is
// positive (x - lower >= 0, a positive divisor, sqrt or t * t of t >= 0, a
// positive factor), so the image of an extreme is the extreme of the images,
// bit for bit.  fast is set only when every range is positive and every
// extreme finite and nonzero (the reductions order -0 below +0, which a
// comparison does not).
__device__ ChainConsts chain_derive(const PSum &so, const PSum &sl, double t_min, double t_max,
                                    double level)
{
    bool ok = isfinite(t_min) && isfinite(t_max) && isfinite(level) && (so.ne || sl.ne);
    bool any = false;
    double mn = 0.0, mx = 0.0;
    auto take = [&](double a, double b) {
        ok = ok && isfinite(a) && isfinite(b) && a != 0.0 && b != 0.0;
        mn = !any || a < mn ? a : mn;
        mx = !any || b > mx ? b : mx;
        any = true;
    };
    ChainConsts c;
    double up1 = t_min;
    c.lo1 = t_max;
    psum_bounds(so, t_min, t_max, c.lo1, up1);  // as k_pr_map, mode 1
    c.r1 = up1 - c.lo1;
    double fo_mn = 0.0, fo_mx = 0.0;
    if (so.ne) {
        ok = ok && c.r1 > 0.0 && isfinite(c.r1);
        fo_mn = pr_power_map(so.mn, c.lo1, c.r1, 0.5);
        fo_mx = pr_power_map(so.mx, c.lo1, c.r1, 0.5);
        take(fo_mn, fo_mx);
    }
    if (sl.ne) take(sl.mn, sl.mx);
    double up0 = mn;
    c.lo0 = mx;
    psum_bounds(sl, mn, mx, c.lo0, up0);  // as k_pr_map, mode 0
    c.r0 = up0 - c.lo0;
    any = false;
    if (so.ne) take(fo_mn, fo_mx);
    if (sl.ne) {
        ok = ok && c.r0 > 0.0 && isfinite(c.r0);
        take(pr_power_map(sl.mn, c.lo0, c.r0, 2.0), pr_power_map(sl.mx, c.lo0, c.r0, 2.0));
    }
    c.h4min = mn - level;
    c.h4max = mx - level;
    ok = ok && isfinite(c.h4min) && isfinite(c.h4max) && c.h4min != 0.0 && c.h4max != 0.0;
    c.fast = ok ? 1 : 0;
    return c;
}

__device__ __forceinline__ double wave_fmin(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_fmax(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double lane_double(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// One half's run within a wave, in sequence order, the same in every lane:
// before its first break the run is strictly decreasing, so its minimum is
// `last` and its maximum `first`; the values from the break on (the
// breaking value included) are taken by the lanes' (mn, mx).
struct WaveRun {
    double first, last;  // last: the last value before the break
    int ne, brk;
};

// Append one slice of 128 values to an unbroken run: lane l holds the
// values at positions 2l (a, if fa) and 2l + 1 (b, if fb) of the slice, the
// values of this half only.  Finds each value's predecessor (the last value
// before it in the slice, else the run's last), the first value not below its
// predecessor (the break), and gives the values from the break on to the
// lane's (mn, mx).
__device__ __forceinline__ void wave_run_slice(WaveRun &R, double a, double b, bool fa, bool fb,
                                               double &mn, double &mx)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long hb = __ballot(fa || fb);
    if (hb == 0) return;
    // Inclusive scan: the last value at or before this lane.
    double pv = fb ? b : a;
    int ph = (fa || fb) ? 1 : 0;
    for (int off = 1; off < 64; off <<= 1) {
        const double t = __shfl_up(pv, off, 64);
        const int th = __shfl_up(ph, off, 64);
        if (lane >= off && !ph) {
            pv = t;
            ph = th;
        }
    }
    // The predecessor of this lane's first value.
    double ep = __shfl_up(pv, 1, 64);
    int eh = __shfl_up(ph, 1, 64);
    if (lane == 0) eh = 0;
    if (!eh) {
        ep = R.last;
        eh = R.ne;
    }
    const bool ba = fa && eh && !(a < ep);
    const double pb = fa ? a : ep;  // the predecessor of b
    const bool bb = fb && (fa || eh) && !(b < pb);
    if (!R.ne) R.first = lane_double(fa ? a : b, __ffsll((long long)hb) - 1);
    R.ne = 1;
    const unsigned long long bm = __ballot(ba || bb);
    if (bm == 0) {
        R.last = lane_double(pv, 63 - __clzll((long long)hb));
        return;
    }
    const int L = __ffsll((long long)bm) - 1;
    const int at_b = __builtin_amdgcn_readlane(ba ? 0 : 1, L);
    R.last = lane_double(at_b ? pb : ep, L);
    R.brk = 1;
    const int P = 2 * L + at_b;  // the break's position
    if (fa && 2 * lane >= P) {
        mn = a < mn ? a : mn;
        mx = a > mx ? a : mx;
    }
    if (fb && 2 * lane + 1 >= P) {
        mn = b < mn ? b : mn;
        mx = b > mx ? b : mx;
    }
}

// The wave's summary of one half (valid in every lane).
__device__ __forceinline__ PSum wave_run_summary(const WaveRun &R, double mn, double mx)
{
    PSum r = psum_empty();
    if (!R.ne) return r;
    r.ne = 1;
    r.brk = R.brk;
    r.first = R.first;
    r.last = R.last;
    r.mn = R.last;
    r.mx = R.first;
    if (R.brk) {
        mn = wave_fmin(mn);
        mx = wave_fmax(mx);
        r.mn = mn < r.mn ? mn : r.mn;
        r.mx = mx > r.mx ? mx : r.mx;
        r.suf = mx;
    }
    return r;
}

// In-order join of the runs held by the block's threads (thread t holds run
// t; empty runs are skipped), in parallel; the result is valid in every
// thread.  Joined in order, the runs stay one strictly decreasing run until
// the first run K that breaks: one broken inside, or one whose first value
// is not below the last value of the nearest non-empty run before it.  So
// first is the first non-empty run's, min and max are over all runs, and
// after a break suf is run K's (its max if the break is at its start, else
// its own suf) joined with the max of every later run; without one, last is
// the last non-empty run's.  A lane scan finds each run's predecessor,
// ballots the first break and the first and last runs.
template <int NT>
__device__ PSum psum_join_block(const PSum &P)
{
    constexpr int NW = NT / 64;
    __shared__ double s_last[NW], s_mn[NW], s_mx[NW], s_sf[NW];
    __shared__ int s_has[NW], s_k[NW], s_f[NW], s_l[NW];
    __shared__ double s_first, s_lastv;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int ne = P.ne;
    // inclusive scan over the wave: the last value of the latest non-empty run
    double pv = P.last;
    int ph = ne;
    for (int off = 1; off < 64; off <<= 1) {
        const double tv = __shfl_up(pv, off, 64);
        const int th = __shfl_up(ph, off, 64);
        if (lane >= off && !ph) {
            pv = tv;
            ph = th;
        }
    }
    if (lane == 63) {
        s_last[w] = pv;
        s_has[w] = ph;
    }
    const unsigned long long bn = __ballot(ne);
    if (lane == 0) {
        s_f[w] = bn ? w * 64 + __ffsll((long long)bn) - 1 : NT;
        s_l[w] = bn ? w * 64 + 63 - __clzll((long long)bn) : -1;
    }
    double mn = ne ? P.mn : INFINITY, mx = ne ? P.mx : -INFINITY;
    mn = wave_fmin(mn);
    mx = wave_fmax(mx);
    if (lane == 0) {
        s_mn[w] = mn;
        s_mx[w] = mx;
    }
    __syncthreads();
    // the predecessor: within the wave, else the latest earlier wave's
    double ep = __shfl_up(pv, 1, 64);
    int eh = __shfl_up(ph, 1, 64);
    if (lane == 0) eh = 0;
    if (!eh) {
        for (int j = w - 1; j >= 0; --j)
            if (s_has[j]) {
                ep = s_last[j];
                eh = 1;
                break;
            }
    }
    const bool cross = ne && eh && !(P.first < ep);
    const bool brk = ne && (P.brk || cross);
    const unsigned long long bb = __ballot(brk);
    if (lane == 0) s_k[w] = bb ? w * 64 + __ffsll((long long)bb) - 1 : NT;
    int F = NT, L = -1;
    double MN = INFINITY, MX = -INFINITY;
    for (int j = 0; j < NW; ++j) {
        F = s_f[j] < F ? s_f[j] : F;
        L = s_l[j] > L ? s_l[j] : L;
        MN = fmin(MN, s_mn[j]);
        MX = fmax(MX, s_mx[j]);
    }
    if (t == F) s_first = P.first;
    if (t == L) s_lastv = P.last;
    __syncthreads();
    int K = NT;
    for (int j = 0; j < NW; ++j) K = s_k[j] < K ? s_k[j] : K;
    double sf = (ne && t > K) ? P.mx : -INFINITY;
    if (t == K) sf = cross ? P.mx : P.suf;
    sf = wave_fmax(sf);
    if (lane == 0) s_sf[w] = sf;
    __syncthreads();
    PSum r = psum_empty();
    if (F < NT) {
        r.ne = 1;
        r.first = s_first;
        r.last = s_lastv;
        r.mn = MN;
        r.mx = MX;
        if (K < NT) {
            double SF = -INFINITY;
            for (int j = 0; j < NW; ++j) SF = fmax(SF, s_sf[j]);
            r.brk = 1;
            r.suf = SF;
        }
    }
    __syncthreads();  // the shared scalars are reused by the next call
    return r;
}

// In-order reduction of the nsum pairs of block summaries sum[2c], sum[2c+1]
// by one block of NT threads: each thread joins a run of ceil(nsum / NT)
// consecutive pairs, then the block joins the threads' results in parallel;
// the results are valid in every thread.
template <int NT>
__device__ void pr_reduce_sums(const PSum *sum, int64_t nsum, PSum &s0, PSum &s1)
{
    const int64_t per = (nsum + NT - 1) / NT;
    const int64_t c0 = (int64_t)threadIdx.x * per;
    const int64_t c1 = c0 + per < nsum ? c0 + per : nsum;
    s0 = psum_empty();
    s1 = psum_empty();
    for (int64_t c = c0; c < c1; ++c) {
        s0 = psum_join(s0, sum[2 * c]);
        s1 = psum_join(s1, sum[2 * c + 1]);
    }
    s0 = psum_join_block<NT>(s0);
    s1 = psum_join_block<NT>(s1);
}

// Summaries of the values whose mask is unset (sum[2b]) and set (sum[2b + 1])
// over block b's values.  Wave w of block b takes the contiguous run of
// wave chunks [(8b + w) * cpw, (8b + w + 1) * cpw), each as four slices of
// 128 values, two per lane.  Each half's run is followed in sequence order
// (WaveRun, wave_run_slice) until its first break (a value not below the
// one before it); after the break only the minimum and maximum of that
// half's later values can change the summary (a later value that is a new
// running minimum is below the maximum already taken), so they are kept
// per lane and reduced once at the end.  A chunk in which no unbroken half
// has a value takes the minimum/maximum path alone: waves whose run lies
// inside one region see a single half for all of it, and every wave must
// stream at the same rate.  The eight waves join in order.
// CHAIN: x holds the noise heights; each value is the standard rescale
// (util.py:139) with min/max from head->keys0, the rescaled min/max and the
// ocean level are derived from the same keys (the map is monotone, so its
// extremes are the images of x's), and mask = t <= level (terrain.py:61-72);
// WRITE: t and the mask are written back.  vec: x is 16-byte and mask
// 2-byte aligned (vector accesses).
template <bool CHAIN, bool WRITE>
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
    WaveRun R0 = {0.0, 0.0, 0, 0}, R1 = {0.0, 0.0, 0, 0};  // land, ocean
    double mn0 = INFINITY, mx0 = -INFINITY, mn1 = INFINITY, mx1 = -INFINITY;  // after the break
    const int64_t u_begin = ((int64_t)blockIdx.x * (kPrThreads / 64) + wave) * cpw;
    const int64_t u_end = u_begin + cpw < nunits ? u_begin + cpw : nunits;
    // Wave chunk u is kPrSlices slices of 128 values; lane l holds values
    // 2l and 2l + 1 of each slice (v[2k], v[2k + 1] of slice k), so every
    // access instruction covers one contiguous range.  Value j of the lane is
    // at p0 + off(j) and exists when off(j) < left = n - p0.
    auto off = [](int j) { return (int64_t)(j >> 1) * 128 + (j & 1); };
    auto load = [&](int64_t u, double *v, unsigned &mb, int64_t &left) {
        const int64_t p0 = u * kPrUnit + 2 * lane;
        left = n - p0;
        mb = 0;
        if (vec && off(kPrPer - 1) < left) {  // the lane's whole chunk
#pragma unroll
            for (int k = 0; k < kPrSlices; ++k) {
                const double2 t = *(const double2 *)(x + p0 + k * 128);
                v[2 * k] = t.x;
                v[2 * k + 1] = t.y;
                if (!CHAIN) {
                    const unsigned w = *(const uint16_t *)(mask + p0 + k * 128);
                    mb |= ((w & 0xffu) != 0 ? 1u : 0u) << (2 * k);
                    mb |= ((w >> 8) != 0 ? 1u : 0u) << (2 * k + 1);
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < kPrPer; ++j) {
                const bool ok = off(j) < left;
                v[j] = ok ? x[p0 + off(j)] : 0.0;
                if (!CHAIN && ok && mask[p0 + off(j)] != 0) mb |= 1u << j;
            }
        }
    };
    double vn[kPrPer];
    unsigned mbn = 0;
    int64_t leftn = 0;
    if (u_begin < u_end) load(u_begin, vn, mbn, leftn);
    for (int64_t u = u_begin; u < u_end; ++u) {
        const int64_t p0 = u * kPrUnit + 2 * lane;
        double v[kPrPer];
        unsigned mb = mbn;  // bit j: value j's mask is set
        const int64_t left = leftn;
#pragma unroll
        for (int k = 0; k < kPrPer; ++k) v[k] = vn[k];
        if (u + 1 < u_end) load(u + 1, vn, mbn, leftn);  // in flight while u is processed
        if (CHAIN) {
            mb = 0;
#pragma unroll
            for (int k = 0; k < kPrPer; ++k) {
                v[k] = ((v[k] - x_min) / xr) * nr + ca.lower;
                mb |= (v[k] <= level ? 1u : 0u) << k;
            }
        }
        if (CHAIN && WRITE) {
            if (vec && off(kPrPer - 1) < left) {
#pragma unroll
                for (int k = 0; k < kPrSlices; ++k) {
                    *(double2 *)(x + p0 + k * 128) = make_double2(v[2 * k], v[2 * k + 1]);
                    *(uint16_t *)(mask + p0 + k * 128) =
                        (uint16_t)(((mb >> (2 * k)) & 1u) | (((mb >> (2 * k + 1)) & 1u) << 8));
                }
            } else {
#pragma unroll
                for (int j = 0; j < kPrPer; ++j)
                    if (off(j) < left) {
                        x[p0 + off(j)] = v[j];
                        mask[p0 + off(j)] = (uint8_t)((mb >> j) & 1u);
                    }
            }
        }
        // bit j of okb: value j exists and is not NaN; ob / lb: of the
        // ocean (mask set) / land half.
        unsigned okb = 0;
#pragma unroll
        for (int k = 0; k < kPrPer; ++k) okb |= (off(k) < left && v[k] == v[k] ? 1u : 0u) << k;
        const unsigned ob = okb & mb, lb = okb & ~mb;
        // A half that has broken, or has no value in this chunk, takes its
        // values' minimum and maximum only.
        if ((R0.brk || !__any(lb != 0)) && (R1.brk || !__any(ob != 0))) {
#pragma unroll
            for (int k = 0; k < kPrPer; ++k) {
                const bool o = (ob >> k) & 1u, l = (lb >> k) & 1u;
                mn1 = o && v[k] < mn1 ? v[k] : mn1;
                mx1 = o && v[k] > mx1 ? v[k] : mx1;
                mn0 = l && v[k] < mn0 ? v[k] : mn0;
                mx0 = l && v[k] > mx0 ? v[k] : mx0;
            }
            continue;
        }
        // Otherwise slice by slice, in order (a rare path, kept small: the
        // slice's two values are selected from v).
#pragma unroll 1
        for (int k = 0; k < kPrSlices; ++k) {
            double a = v[0], b = v[1];
#pragma unroll
            for (int t = 1; t < kPrSlices; ++t) {
                a = k == t ? v[2 * t] : a;
                b = k == t ? v[2 * t + 1] : b;
            }
            const bool la = (lb >> (2 * k)) & 1u, lbb = (lb >> (2 * k + 1)) & 1u;
            const bool oa = (ob >> (2 * k)) & 1u, obb = (ob >> (2 * k + 1)) & 1u;
            if (R0.brk) {
                mn0 = la && a < mn0 ? a : mn0;
                mx0 = la && a > mx0 ? a : mx0;
                mn0 = lbb && b < mn0 ? b : mn0;
                mx0 = lbb && b > mx0 ? b : mx0;
            } else {
                wave_run_slice(R0, a, b, la, lbb, mn0, mx0);
            }
            if (R1.brk) {
                mn1 = oa && a < mn1 ? a : mn1;
                mx1 = oa && a > mx1 ? a : mx1;
                mn1 = obb && b < mn1 ? b : mn1;
                mx1 = obb && b > mx1 ? b : mx1;
            } else {
                wave_run_slice(R1, a, b, oa, obb, mn1, mx1);
            }
        }
    }
    const PSum w0 = wave_run_summary(R0, mn0, mx0), w1 = wave_run_summary(R1, mn1, mx1);
    if (lane == 0) {
        lds[0][wave] = w0;
        lds[1][wave] = w1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        PSum r0 = lds[0][0], r1 = lds[1][0];
        for (int w = 1; w < kPrThreads / 64; ++w) {
            r0 = psum_join(r0, lds[0][w]);
            r1 = psum_join(r1, lds[1][w]);
        }
        sum[2 * blockIdx.x] = r0;
        sum[2 * blockIdx.x + 1] = r1;
    }
}

// One block: in-order reduction of the per-block summaries into head->s;
// from_keys: head->xmin/xmax = (min, max) of x from head->keys0; chain: the
// two-pass height chain's constants (chain_derive) into head, and the
// five-pass kernels' reduction keys reset for the case they run.
__global__ __launch_bounds__(1024) void k_pr_finish(const PSum *__restrict__ sum, int64_t nsum,
                                                    PrHead *__restrict__ head, int from_keys,
                                                    int chain)
{
    PSum s0, s1;
    pr_reduce_sums<1024>(sum, nsum, s0, s1);
    if (threadIdx.x == 0) {
        head->s[0] = s0;
        head->s[1] = s1;
        if (from_keys) {
            head->xmin = dunkey(head->keys0[0]);
            head->xmax = dunkey(head->keys0[1]);
        }
        if (chain) {
            const ChainConsts cc = chain_derive(s1, s0, head->xmin, head->xmax, head->level);
            head->lo1 = cc.lo1;
            head->r1 = cc.r1;
            head->lo0 = cc.lo0;
            head->r0 = cc.r0;
            head->h4min = cc.h4min;
            head->h4max = cc.h4max;
            head->fast = cc.fast;
            head->keys1[0] = head->keys2[0] = ~0ull;
            head->keys1[1] = head->keys2[1] = 0ull;
        }
    }
}

__device__ __forceinline__ bool pr_sel(int mode, uint8_t m) { return mode == 1 ? m != 0 : (mode == 0 && m == 0); }

// util.py:224-248 for the selected values; the others are copied.  The
// bounds start from (min, max) of the input: from keys when given, else from
// head->xmin/xmax.  shift: subtract head->level from every value afterwards
// (nixis.py:362).  okeys (nullable): min/max keys of the values written.
// x and out may alias.  vec: x and out 16-byte, mask 2-byte aligned.
// skip_if_fast: return at once when head->fast (the two-pass chain wrote
// the results).
__global__ __launch_bounds__(256) void k_pr_map(const double *x, double *out,
                                                const uint8_t *__restrict__ mask, int64_t n,
                                                int mode, double power,
                                                const PrHead *__restrict__ head,
                                                const unsigned long long *__restrict__ keys,
                                                int shift, unsigned long long *__restrict__ okeys,
                                                int *__restrict__ err, int vec, int skip_if_fast)
{
    if (skip_if_fast && head->fast) return;
    const double x_min = keys ? dunkey(keys[0]) : head->xmin;
    const double x_max = keys ? dunkey(keys[1]) : head->xmax;
    double mask_lower = x_max, mask_upper = x_min;
    if (mode == 0 || mode == 1) psum_bounds(head->s[mode], x_min, x_max, mask_lower, mask_upper);
    const double level = shift ? head->level : 0.0;
    const double mask_range = mask_upper - mask_lower;
    bool zero = false;
    unsigned long long lo = ~0ull, hi = 0ull;
    auto f = [&](double r, uint8_t m) {
        if (pr_sel(mode, m)) {
            zero = zero | (mask_range == 0.0);
            r = pr_power_map(r, mask_lower, mask_range, power);
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

// The height chain's second pass (nixis.py:332-364 per value), after
// k_pr_finish has derived the chain's constants into head: the first
// rescale with min/max from head->keys0 and the mask t <= level; when the
// constants hold (head->fast), the ocean or land power_rescale map, the
// shift by the level and the split rescale around 0 with the shifted
// heights' min/max.  Heights and mask are written once.  Every operation is
// the one the five-pass kernels apply to that value, in the same order.
// vec: h 16-byte and mask 2-byte aligned.
__global__ __launch_bounds__(kFinalThreads) void k_chain_final(double *h, uint8_t *__restrict__ mask,
                                                               int64_t n,
                                                               const PrHead *__restrict__ head,
                                                               double min_alt, double max_alt,
                                                               int *__restrict__ err, int vec)
{
    const double x_min = dunkey(head->keys0[0]), x_max = dunkey(head->keys0[1]);
    const double xr = x_max - x_min, nr = max_alt - min_alt;
    const double level = head->level;
    const double lo1 = head->lo1, r1 = head->r1, lo0 = head->lo0, r0 = head->r0;
    const double h4min = head->h4min, h4max = head->h4max;
    const bool fast = head->fast != 0;
    bool zero = false;
    auto f = [&](double v, uint8_t &m) {
        const double t = ((v - x_min) / xr) * nr + min_alt;
        m = t <= level ? 1 : 0;
        if (!fast) return t;  // the five-pass kernels continue from here
        double r = m ? pr_power_map(t, lo1, r1, 0.5) : pr_power_map(t, lo0, r0, 2.0);
        r = r - level;
        return split_rescale(r, h4min, h4max, min_alt, max_alt, zero);
    };
    if (vec) {
        double2 *h2 = (double2 *)h;
        uint16_t *m2 = (uint16_t *)mask;
        stream4<double2>(n >> 1, [&](int64_t j) { return h2[j]; }, [&](int64_t j, double2 d) {
            uint8_t ma, mb;
            const double a = f(d.x, ma), b = f(d.y, mb);
            h2[j] = make_double2(a, b);
            m2[j] = (uint16_t)(ma | (mb << 8));
        });
        if ((n & 1) && tail_thread()) {
            uint8_t m;
            h[n - 1] = f(h[n - 1], m);
            mask[n - 1] = m;
        }
    } else {
        stream4<double>(n, [&](int64_t i) { return h[i]; }, [&](int64_t i, double v) {
            uint8_t m;
            h[i] = f(v, m);
            mask[i] = m;
        });
    }
    // chain_derive sets fast only when no range is zero; kept as a check.
    if (__any(zero) && (threadIdx.x & 63) == 0)
        __hip_atomic_fetch_or(err, NX_ERR_DIV0_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

// Summary blocks: at most kPrMaxBlocks; each wave takes a contiguous run of
// wave chunks.
constexpr int64_t kPrMaxBlocks = 2048;

struct PrGrid {
    int64_t nunits, cpw, blocks;
};

// Resident blocks per CU of one summary kernel (its register use sets it):
// the grid is one round of resident blocks, every wave a contiguous run.
template <bool CHAIN, bool WRITE>
static int pr_blocks_per_cu()
{
    static std::atomic<int> cached{0};
    int v = cached.load(std::memory_order_relaxed);
    if (v == 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, reinterpret_cast<const void *>(&k_pr_summary<CHAIN, WRITE>), kPrThreads, 0) !=
                hipSuccess ||
            nb < 1)
            nb = 1;
        v = nb;
        cached.store(v, std::memory_order_relaxed);
    }
    return v;
}

static PrGrid pr_grid(nx_ctx *ctx, int64_t n, int blocks_per_cu)
{
    constexpr int64_t wpb = kPrThreads / 64;
    PrGrid g;
    g.nunits = (n + kPrUnit - 1) / kPrUnit;
    const int64_t want = wpb * std::max<int64_t>(1, std::min<int64_t>(kPrMaxBlocks, (int64_t)ctx->num_cus * blocks_per_cu));
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
    return ((uintptr_t)x % 16 == 0 && (uintptr_t)mask % 2 == 0) ? 1 : 0;
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
    const PrGrid g = pr_grid(ctx, n, pr_blocks_per_cu<false, false>());
    const unsigned grid = red_grid(ctx, n);
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, stream, head->keys0);
    hipLaunchKernelGGL(k_minmax, dim3(mm_grid(ctx, n)), dim3(kMinmaxThreads), 0, stream, x, n, head->keys0,
                       vec16(x));
    hipLaunchKernelGGL((k_pr_summary<false, false>), dim3((unsigned)g.blocks), dim3(kPrThreads), 0,
                       stream, const_cast<double *>(x), const_cast<uint8_t *>(mask), n, g.nunits,
                       g.cpw, pr_vec(x, mask), sum, head, ChainArgs{0.0, 0.0, 0.0, nullptr},
                       ctx->d_err);
    hipLaunchKernelGGL(k_pr_finish, dim3(1), dim3(1024), 0, stream, sum, g.blocks, head, 1, 0);
    hipLaunchKernelGGL(k_pr_map, dim3(grid), dim3(256), 0, stream, x, out, mask, n, mode, power,
                       head, (const unsigned long long *)nullptr, 0,
                       (unsigned long long *)nullptr, ctx->d_err, vec16(x) & vec16(out) & vec2(mask), 0);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

// ---- nixis.py:330-364: the chain after noise ------------------------------------
//
// Two passes over the heights after their min/max: the first computes the
// rescale and the mask of each value and the power_rescale summaries of both
// halves without writing (the land half's values are not changed by the
// ocean half's power_rescale, so one pass summarises both); one block joins
// the block summaries and derives every later bound and min/max from them
// (k_pr_finish, chain_derive); the second pass applies the whole chain to
// each value and writes heights and mask.  When the derivation does not hold (a zero range,
// a zero or non-finite extreme), the second pass writes the rescaled heights
// and mask instead, and the kernels of the five-pass form, which otherwise
// return at once, finish the chain: the ocean power_rescale with the min/max
// of its result; the land power_rescale with the shift by the ocean level and
// the min/max of its result; the final split rescale around 0.  The
// five-pass form itself (NX_OPT_CHAIN_PASSES 5) writes in its summary pass
// and reduces the summaries in a kernel of its own.  Both forms give the same
// bits.

size_t height_chain_workspace_bytes(int64_t n) { return power_rescale_workspace_bytes(n); }

// The chain after head->keys0 holds the min/max keys of h.
static int chain_after_keys(nx_ctx *ctx, double *h, int64_t n, double min_alt, double max_alt,
                            double ocean_percent, uint8_t *mask, double *level, void *ws,
                            hipStream_t stream)
{
    if (!(0.0 < ocean_percent && ocean_percent < 100.0)) ocean_percent = 50.0;  // util.py:561-564
    auto *head = (PrHead *)ws;
    auto *sum = (PSum *)((char *)ws + 256);
    const int two = ctx->chain_passes == 2 ? 1 : 0;
    const PrGrid g = pr_grid(ctx, n, two ? pr_blocks_per_cu<true, false>() : pr_blocks_per_cu<true, true>());
    const unsigned grid = red_grid(ctx, n);
    const ChainArgs ca{min_alt, max_alt, ocean_percent, level};
    const int vsum = pr_vec(h, mask), vmap = vec16(h) & vec2(mask);
    if (ctx->chain_passes == 2) {
        // The summary pass, the join and derivation, then the writing pass;
        // the five-pass kernels after it return at once when head->fast,
        // else continue from the rescaled heights and mask the writing pass
        // left (k_pr_finish resets keys1 and keys2).
        hipLaunchKernelGGL((k_pr_summary<true, false>), dim3((unsigned)g.blocks), dim3(kPrThreads),
                           0, stream, h, mask, n, g.nunits, g.cpw, vsum, sum, head, ca, ctx->d_err);
        hipLaunchKernelGGL(k_pr_finish, dim3(1), dim3(1024), 0, stream, sum, g.blocks, head, 0, 1);
        hipLaunchKernelGGL(k_chain_final, dim3(grid), dim3(kFinalThreads), 0, stream, h, mask, n,
                           (const PrHead *)head, min_alt, max_alt, ctx->d_err, vmap);
    } else {
        // nixis.py:332-346 in one pass
        hipLaunchKernelGGL((k_pr_summary<true, true>), dim3((unsigned)g.blocks),
                           dim3(kPrThreads), 0, stream, h, mask, n, g.nunits, g.cpw, vsum, sum, head,
                           ca, ctx->d_err);
        hipLaunchKernelGGL(k_pr_finish, dim3(1), dim3(1024), 0, stream, sum, g.blocks, head, 0, 0);
        hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, stream, head->keys1);
        hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, stream, head->keys2);
    }
    // nixis.py:352: the ocean floor, power 0.5
    hipLaunchKernelGGL(k_pr_map, dim3(grid), dim3(256), 0, stream, h, h, mask, n, 1, 0.5, head,
                       (const unsigned long long *)nullptr, 0, head->keys1, ctx->d_err, vmap, two);
    // nixis.py:355 and 362: the land, power 2.0, then the shift
    hipLaunchKernelGGL(k_pr_map, dim3(grid), dim3(256), 0, stream, h, h, mask, n, 0, 2.0, head,
                       (const unsigned long long *)head->keys1, 1, head->keys2, ctx->d_err, vmap,
                       two);
    // nixis.py:364
    RescaleArgs mid_args{min_alt, max_alt, 0.0, 0.0, 0.0, 1, 0, 0, 0};
    hipLaunchKernelGGL(k_rescale_ex, dim3(grid), dim3(256), 0, stream, h, n,
                       (const unsigned long long *)head->keys2, mid_args, h, (double *)nullptr,
                       ctx->d_err, vec16(h), two ? (const int32_t *)&head->fast : (const int32_t *)nullptr);
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
    hipLaunchKernelGGL(k_minmax, dim3(mm_grid(ctx, n)), dim3(kMinmaxThreads), 0, stream, h, n, head->keys0,
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
