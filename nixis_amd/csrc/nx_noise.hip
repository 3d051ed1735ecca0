// nx_noise.hip -- fused multi-octave OpenSimplex noise over icosphere vertices.
//
// terrain.py:32-59 (sample_octaves) calls terrain.py:12-28 (sample_noise)
// once per octave; each call rescales all coordinates, evaluates noise3d at
// every vertex and adds ((e + 1) * 0.5 * strength) * radius to the
// elevations.  Here one lane owns one vertex for all octaves: coordinates
// are read once, each octave's coordinate scale, skew and lattice selection
// stay in FP64 in the reference's order (or, when the frequencies double,
// the skewed coordinates are doubled from the previous octave, which gives
// the same bits), the octaves' weighted FP32 sums accumulate in FP32 and
// join the FP64 sum of the octave offsets, and only the final elevation is
// written.  Tables live in LDS (see nx_noise_dev.h).
#include <hip/hip_runtime.h>

#include <cstring>

#include "nx_ctx.h"
#include "nx_noise_dev.h"
#include "nx_reduce_dev.h"

namespace nx {

struct OctaveArgs {
    NoiseTablesPacked tables;
    OctavePlan oct;
    int accumulate;
};

// keys (nullable): order-preserving min/max keys of the elevations written,
// reduced here so a rescale after noise needs no pass of its own.
template <bool DBG>
__device__ __forceinline__ void sample_octaves_body(const OctaveArgs &a,
                                                    const double *__restrict__ verts, int64_t V,
                                                    double *__restrict__ elev,
                                                    float *__restrict__ dbg_noise,
                                                    int32_t *__restrict__ dbg_sel,
                                                    int *__restrict__ err,
                                                    unsigned long long *__restrict__ keys)
{
    unsigned long long lo = ~0ull, hi = 0ull;
    __shared__ NoiseTables T;
    load_tables(T, a.tables);
    __syncthreads();
    bool bad = false;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += stride) {
        const double x = verts[3 * i + 0], y = verts[3 * i + 1], z = verts[3 * i + 2];
        const double acc0 = a.accumulate ? elev[i] : 0.0;
        const double acc = octave_sum<DBG>(x, y, z, acc0, a.oct, T, bad,
                                           [&](int o, float v, const NoiseSel &sel) {
            if (DBG) {
                if (dbg_noise) dbg_noise[(int64_t)o * V + i] = v * (kSumScale / 103.0f);
                if (dbg_sel) {
                    int32_t *d = dbg_sel + ((int64_t)o * V + i) * 5;
                    for (int k = 0; k < 5; ++k) d[k] = sel.v[k];
                }
            }
        });
        elev[i] = acc;
        if (keys) {
            const unsigned long long k = dkey(acc);
            lo = k < lo ? k : lo;
            hi = k > hi ? k : hi;
        }
    }
    if (bad) __hip_atomic_fetch_or(err, NX_ERR_RANGE_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (keys) {
        block_minmax(lo, hi);
        if (threadIdx.x == 0) {
            __hip_atomic_fetch_min(&keys[0], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_max(&keys[1], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// The VALU-bound launch asks for 8 waves per SIMD (at most 64 VGPRs), which
// keeps enough waves in flight to cover the dependent LDS hash lookups.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_sample_octaves(
    OctaveArgs a, const double *__restrict__ verts, int64_t V, double *__restrict__ elev,
    int *__restrict__ err, unsigned long long *__restrict__ keys)
{
    sample_octaves_body<false>(a, verts, V, elev, nullptr, nullptr, err, keys);
}

// Debug launch: per-octave noise values and selection records.
__global__ __launch_bounds__(256) void k_sample_octaves_dbg(
    OctaveArgs a, const double *__restrict__ verts, int64_t V, double *__restrict__ elev,
    float *__restrict__ dbg_noise, int32_t *__restrict__ dbg_sel, int *__restrict__ err)
{
    sample_octaves_body<true>(a, verts, V, elev, dbg_noise, dbg_sel, err, nullptr);
}

__global__ __launch_bounds__(256) void k_noise3d(NoiseTablesPacked tp, const double *__restrict__ x,
                                                 const double *__restrict__ y,
                                                 const double *__restrict__ z, int64_t n,
                                                 double *__restrict__ out, int *__restrict__ err)
{
    __shared__ NoiseTables T;
    load_tables(T, tp);
    __syncthreads();
    bool bad = false;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (double)noise3<false>(x[i], y[i], z[i], T, bad, nullptr);
    if (bad) __hip_atomic_fetch_or(err, NX_ERR_RANGE_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Pack perm / pgi (host int32[256] each) into the kernel-argument form.  The
// reference indexes both tables without bounds checks (opensimplex.py:125);
// values outside 0..255 (perm) or outside the 24 gradient offsets (pgi)
// would read past its arrays, so they are rejected here.
int pack_tables(const int32_t *perm, const int32_t *pgi, NoiseTablesPacked &p)
{
    if (!perm || !pgi) return set_error(NX_EINVAL, "perm/pgi must not be NULL");
    uint8_t pb[256], gb[256];
    for (int i = 0; i < 256; ++i) {
        if (perm[i] < 0 || perm[i] > 255)
            return set_error(NX_EINVAL, "perm[%d]=%d outside 0..255", i, perm[i]);
        if (pgi[i] < 0 || pgi[i] > 69 || pgi[i] % 3 != 0)
            return set_error(NX_EINVAL, "pgi[%d]=%d is not a gradient index", i, pgi[i]);
        pb[i] = (uint8_t)perm[i];
        gb[i] = (uint8_t)(8 * (pgi[i] / 3));  // byte offset of the gradient's word
    }
    std::memcpy(p.perm, pb, 256);
    std::memcpy(p.grad, gb, 256);
    // opensimplex.py:51-60: gradient g = 3s + k has magnitude 11 on axis k and
    // 4 on the others; the x sign is + for odd s, y and z signs are - when
    // bit 1 / bit 2 of s is set.  Stored as fp16 (exact for +-4, +-11).
    for (int g = 0; g < 24; ++g) {
        const int s = g / 3, k = g % 3;
        const int gx = (k == 0 ? 11 : 4) * ((s & 1) ? 1 : -1);
        const int gy = (k == 1 ? 11 : 4) * ((s & 2) ? -1 : 1);
        const int gz = (k == 2 ? 11 : 4) * ((s & 4) ? -1 : 1);
        p.g16[2 * g] = half_of_small(gx) | (half_of_small(gy) << 16);
        const float fz = (float)gz;
        std::memcpy(&p.g16[2 * g + 1], &fz, 4);
    }
    return NX_OK;
}

// Per-octave frequency and weights (OctavePlan), the octave with the largest
// |frequency| (its domain check bounds the others) and whether the
// frequencies double exactly from one octave to the next (terrain.py:44,
// lacunarity 2).
int octave_plan(const double *rough, const double *strength, double radius, int n_oct,
                OctavePlan &p)
{
    std::memset(&p, 0, sizeof(p));
    double best = -1.0;
    p.doubling = 1;
    for (int o = 0; o < n_oct; ++o) {
        if (!std::isfinite(rough[o]))
            return set_error(NX_EINVAL, "octave %d frequency is not finite", o);
        const double w = (0.5 * strength[o]) * radius;
        p.rough[o] = rough[o];
        p.wsumf[o] = (float)(w * kSumScaleD / 103.0);
        p.base += w;
        if (std::fabs(rough[o]) > best) {
            best = std::fabs(rough[o]);
            p.check_oct = o;
        }
        if (o > 0 && !(rough[o] == 2.0 * rough[o - 1])) p.doubling = 0;
    }
    p.n_oct = n_oct;
    return NX_OK;
}

static int grid_for(nx_ctx *ctx, int64_t n, int per_cu)
{
    const int64_t blocks = (n + 255) / 256;
    const int64_t cap = (int64_t)ctx->num_cus * per_cu;
    return (int)std::max<int64_t>(1, std::min(blocks, cap));
}

int sample_octaves(nx_ctx *ctx, const double *verts, int64_t V, const int32_t *perm,
                   const int32_t *pgi, int n_oct, const double *rough, const double *strength,
                   double radius, double *elev, int accumulate, float *dbg_noise,
                   int32_t *dbg_sel, hipStream_t stream)
{
    return sample_octaves_keys(ctx, verts, V, perm, pgi, n_oct, rough, strength, radius, elev,
                               accumulate, dbg_noise, dbg_sel, nullptr, stream);
}

int sample_octaves_keys(nx_ctx *ctx, const double *verts, int64_t V, const int32_t *perm,
                        const int32_t *pgi, int n_oct, const double *rough,
                        const double *strength, double radius, double *elev, int accumulate,
                        float *dbg_noise, int32_t *dbg_sel, unsigned long long *keys,
                        hipStream_t stream)
{
    if (!ctx) return set_error(NX_EINVAL, "null context");
    if (V < 0 || (V > 0 && (!verts || !elev))) return set_error(NX_EINVAL, "bad verts/elev");
    if (n_oct < 0 || n_oct > NX_MAX_OCTAVES)
        return set_error(NX_EINVAL, "n_octaves=%d outside 0..%d", n_oct, NX_MAX_OCTAVES);
    if (n_oct > 0 && (!rough || !strength)) return set_error(NX_EINVAL, "null octave scalars");
    OctaveArgs a;
    std::memset(&a, 0, sizeof(a));
    int rc = pack_tables(perm, pgi, a.tables);
    if (rc) return rc;
    rc = octave_plan(rough, strength, radius, n_oct, a.oct);
    if (rc) return rc;
    a.accumulate = accumulate ? 1 : 0;
    if (V == 0) return NX_OK;
    NX_HIP(hipSetDevice(ctx->device));
    const int grid = grid_for(ctx, V, 8);
    if (dbg_noise || dbg_sel) {
        if (keys) return set_error(NX_EINVAL, "min/max keys are not reduced with debug output");
        hipLaunchKernelGGL(k_sample_octaves_dbg, dim3(grid), dim3(256), 0, stream, a, verts, V,
                           elev, dbg_noise, dbg_sel, ctx->d_err);
    } else {
        hipLaunchKernelGGL(k_sample_octaves, dim3(grid), dim3(256), 0, stream, a, verts, V, elev,
                           ctx->d_err, keys);
    }
    NX_HIP(hipGetLastError());
    return NX_OK;
}

int noise3d(nx_ctx *ctx, const double *x, const double *y, const double *z, int64_t n,
            const int32_t *perm, const int32_t *pgi, double *out, hipStream_t stream)
{
    if (!ctx) return set_error(NX_EINVAL, "null context");
    if (n < 0 || (n > 0 && (!x || !y || !z || !out))) return set_error(NX_EINVAL, "bad arrays");
    NoiseTablesPacked tp;
    int rc = pack_tables(perm, pgi, tp);
    if (rc) return rc;
    if (n == 0) return NX_OK;
    NX_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_noise3d, dim3(grid_for(ctx, n, 8)), dim3(256), 0, stream, tp, x, y, z, n,
                       out, ctx->d_err);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

}  // namespace nx
