// nx_climate.hip -- insolation at the top of the atmosphere on the device
// (climate.py:193-201, 416-590).
//
// Per vertex and rotation step: rotate about z, tilt about y, convert to
// latitude / longitude (util.py:60-76) and add max(cos|lat|, 0) *
// max(cos lon, 0) to a float32 accumulator the way numba adds a double to a
// float32 array element (sum in double, round to float).  The sines and
// cosines of the rotation, tilt and seasonal angles are the same for every
// vertex; they are computed on the host with the C library the reference's
// numba code calls (one sincos per angle, as LLVM pairs a sin and a cos of
// one operand) and uploaded as tables.  The per-vertex asin, atan2 and cos
// run on the device.
#include <hip/hip_runtime.h>

#include <cmath>

#include "nx_ctx.h"

namespace nx {

namespace {

constexpr int kMaxSteps = 360;  // rotation steps per call (a day of 1-degree steps)
constexpr int kDays = 360;      // climate.py:581
constexpr int kSlice = 181;     // latitudes -90..90, index = lat mod 181

struct ClimateLayout {
    size_t rot, tilt, sverts, tables, total;
};

constexpr size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr ClimateLayout climate_layout()
{
    ClimateLayout L{};
    size_t o = 0;
    L.rot = o;
    o += a256(16 * kMaxSteps);       // cos, sin per rotation step
    L.tilt = o;
    o += a256(16 * kDays);           // cos, sin per day's tilt
    L.sverts = o;
    o += a256(24 * kSlice);          // the slice's vertices
    L.tables = o;
    o += a256(4 * (size_t)kDays * kSlice);
    L.total = o;
    return L;
}

__device__ __forceinline__ double py_max(double a, double b) { return b > a ? b : a; }
__device__ __forceinline__ double py_min(double a, double b) { return b < a ? b : a; }

// util.py:60-76: degrees(v) = v * (180 / pi).
__device__ __forceinline__ double lat_of(double z, double r)
{
    return asin(py_min(py_max(z / r, -1.0), 1.0)) * (180.0 / 3.141592653589793);
}

// climate.py:421-440 and 444 for one rotation step.
__device__ __forceinline__ double insol_value(double x, double y, double z, double r, double cr,
                                              double sr, double ct, double st)
{
    const double pi = 3.141592653589793;
    const double rx = x * cr - y * sr;
    const double ry = x * sr + y * cr;
    const double rz = z;
    const double tx = rx * ct + rz * st;
    const double ty = ry;
    const double tz = rz * ct - rx * st;
    const double lat = lat_of(tz, r);
    const double lon = atan2(ty, tx) * (180.0 / 3.141592653589793);
    return py_max(cos(fabs(lat) * pi / 180), 0.0) * py_max(cos(lon * pi / 180), 0.0);
}

// climate.py:351-412 (assign_surface_temp).  The reference's altitude weight
// is 0 (climate.py:373-374), so h2 = rescale(altitudes, 0, 0) is 0 at every
// vertex and every vertex takes the else branch (climate.py:408):
// max(cos(|lat - tilt| * pi / 180), 0), stored as float32.
__global__ __launch_bounds__(256) void k_surface_temp(const double *__restrict__ verts,
                                                      int64_t V, double r, double tilt,
                                                      float *__restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double pi = 3.141592653589793;
    const double dist = lat_of(verts[3 * v + 2], r) - tilt;
    out[v] = (float)py_max(cos(fabs(dist) * pi / 180), 0.0);
}

}  // namespace

// nsteps rotations of each vertex, accumulated into arr (float32).
__global__ __launch_bounds__(256) void k_insolation(const double *__restrict__ verts, int64_t V,
                                                    double radius, const double *__restrict__ rot,
                                                    int nsteps, double ct, double st,
                                                    float *__restrict__ arr)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double x = verts[3 * v], y = verts[3 * v + 1], z = verts[3 * v + 2];
    float acc = arr[v];
    for (int s = 0; s < nsteps; ++s)
        acc = (float)((double)acc + insol_value(x, y, z, radius, rot[2 * s], rot[2 * s + 1], ct, st));
    arr[v] = acc;
}

// climate.py:503-533 for `days` tilts at once: tables[d][k] over the 181
// slice vertices, 360 rotation steps each.
__global__ __launch_bounds__(256) void k_slices(const double *__restrict__ sverts, double radius,
                                                const double *__restrict__ rot, int nsteps,
                                                const double *__restrict__ tilt, int days,
                                                float *__restrict__ tables)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= days * kSlice) return;
    const int d = t / kSlice, k = t - d * kSlice;
    const double x = sverts[3 * k], y = sverts[3 * k + 1], z = sverts[3 * k + 2];
    const double ct = tilt[2 * d], st = tilt[2 * d + 1];
    float acc = 0.0f;
    for (int s = 0; s < nsteps; ++s)
        acc = (float)((double)acc + insol_value(x, y, z, radius, rot[2 * s], rot[2 * s + 1], ct, st));
    tables[t] = acc;
}

// climate.py:547-573 with each of `days` tables; with days > 1 the daily
// values are summed in float32 in day order (climate.py:581-584).
__global__ __launch_bounds__(256) void k_interpolate(const double *__restrict__ verts, int64_t V,
                                                     double radius,
                                                     const float *__restrict__ tables, int days,
                                                     float *__restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double lat = lat_of(verts[3 * v + 2], radius);
    const int64_t lower = (int64_t)floor(lat), upper = (int64_t)ceil(lat);
    const int il = (int)(lower < 0 ? kSlice + lower : lower);
    const int iu = (int)(upper < 0 ? kSlice + upper : upper);
    const double frac = lat - (double)lower;
    const double span = (double)(upper - lower);
    float acc = 0.0f;
    for (int d = 0; d < days; ++d) {
        const float tl = tables[d * kSlice + il], tu = tables[d * kSlice + iu];
        float val;
        if (lower == upper) {
            val = tl;
        } else {
            const float diff = tu - tl;
            val = (float)((double)tl + frac * ((double)diff / span));
        }
        acc = days == 1 ? val : acc + val;
    }
    out[v] = acc;
}

size_t climate_workspace_bytes() { return climate_layout().total; }

// rot[2s], rot[2s+1] = cos, sin of (rotation * pi) / 180 for s = 0..nsteps-1.
static void fill_rot(double *t, double rot0, double rot_step, int nsteps)
{
    const double pi = 3.141592653589793;
    double rotation = rot0;
    for (int s = 0; s < nsteps; ++s) {
        sincos(rotation * pi / 180, &t[2 * s + 1], &t[2 * s]);
        rotation += rot_step;
    }
}

static int climate_check(nx_ctx *ctx, const void *ws)
{
    if (!ctx || !ws) return set_error(NX_EINVAL, "insolation needs a context and a workspace");
    if (((uintptr_t)ws & 255) != 0) return set_error(NX_EINVAL, "workspace must be 256-byte aligned");
    return NX_OK;
}

int sample_insolation(nx_ctx *ctx, const double *verts, int64_t V, double radius, double rot0,
                      double rot_step, int nsteps, double tilt, float *arr, void *ws,
                      hipStream_t stream)
{
    int rc = climate_check(ctx, ws);
    if (rc) return rc;
    if (V < 0 || (V > 0 && (!verts || !arr))) return set_error(NX_EINVAL, "bad vertices / output");
    if (nsteps < 0 || nsteps > kMaxSteps)
        return set_error(NX_EINVAL, "nsteps=%d outside 0..%d", nsteps, kMaxSteps);
    if (V == 0 || nsteps == 0) return NX_OK;
    NX_HIP(hipSetDevice(ctx->device));
    const ClimateLayout L = climate_layout();
    int slot = 0;
    double *t = (double *)stage_begin(ctx, 16 * (size_t)nsteps, &slot);
    if (!t) return set_error(NX_ENOMEM, "pinned staging buffer allocation failed");
    fill_rot(t, rot0, rot_step, nsteps);
    if ((rc = stage_commit(ctx, slot, (char *)ws + L.rot, 16 * (size_t)nsteps, stream))) return rc;
    double st, ct;
    sincos(tilt * 3.141592653589793 / 180, &st, &ct);
    hipLaunchKernelGGL(k_insolation, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, stream, verts,
                       V, radius, (const double *)((char *)ws + L.rot), nsteps, ct, st, arr);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

// Slices for `days` tilts into the workspace's tables (climate.py:510-528).
static int slices(nx_ctx *ctx, double radius, const double *tilts, int days, void *ws,
                  hipStream_t stream)
{
    const ClimateLayout L = climate_layout();
    const size_t bytes = L.tables;  // rot, tilt and slice vertices
    int slot = 0;
    char *h = stage_begin(ctx, bytes, &slot);
    if (!h) return set_error(NX_ENOMEM, "pinned staging buffer allocation failed");
    fill_rot((double *)(h + L.rot), -180.0, 360.0 / 360, kMaxSteps);
    double *tl = (double *)(h + L.tilt);
    for (int d = 0; d < days; ++d) sincos(tilts[d] * 3.141592653589793 / 180, &tl[2 * d + 1], &tl[2 * d]);
    double *sv = (double *)(h + L.sverts);
    const double d2r = 3.141592653589793 / 180.0;
    for (int i = -90; i <= 90; ++i) {  // latlon2xyz(i, 0, radius), util.py:86-88
        const int k = i < 0 ? kSlice + i : i;
        double s, c;
        sincos((double)i * d2r, &s, &c);
        sv[3 * k] = radius * c * 1.0;
        sv[3 * k + 1] = radius * c * 0.0;
        sv[3 * k + 2] = radius * s;
    }
    int rc = stage_commit(ctx, slot, ws, bytes, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_slices, dim3((unsigned)((days * kSlice + 255) / 256)), dim3(256), 0, stream,
                       (const double *)((char *)ws + L.sverts), radius,
                       (const double *)((char *)ws + L.rot), kMaxSteps,
                       (const double *)((char *)ws + L.tilt), days, (float *)((char *)ws + L.tables));
    NX_HIP(hipGetLastError());
    return NX_OK;
}

int insolation_slice(nx_ctx *ctx, double radius, double tilt, float *out, void *ws,
                     hipStream_t stream)
{
    int rc = climate_check(ctx, ws);
    if (rc) return rc;
    if (!out) return set_error(NX_EINVAL, "null output");
    NX_HIP(hipSetDevice(ctx->device));
    if ((rc = slices(ctx, radius, &tilt, 1, ws, stream))) return rc;
    NX_HIP(hipMemcpyAsync(out, (char *)ws + climate_layout().tables, sizeof(float) * kSlice,
                          hipMemcpyDeviceToDevice, stream));
    return NX_OK;
}

int daily_insolation(nx_ctx *ctx, const double *verts, int64_t V, double radius, double tilt,
                     float *out, void *ws, hipStream_t stream)
{
    int rc = climate_check(ctx, ws);
    if (rc) return rc;
    if (V < 0 || (V > 0 && (!verts || !out))) return set_error(NX_EINVAL, "bad vertices / output");
    NX_HIP(hipSetDevice(ctx->device));
    if ((rc = slices(ctx, radius, &tilt, 1, ws, stream))) return rc;
    if (V == 0) return NX_OK;
    hipLaunchKernelGGL(k_interpolate, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, stream,
                       verts, V, radius, (const float *)((char *)ws + climate_layout().tables), 1,
                       out);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

int interpolate_insolation(nx_ctx *ctx, const double *verts, int64_t V, const float *table,
                           double radius, float *out, hipStream_t stream)
{
    if (!ctx || V < 0 || (V > 0 && (!verts || !table || !out)))
        return set_error(NX_EINVAL, "bad vertices / table / output");
    if (V == 0) return NX_OK;
    NX_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_interpolate, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, stream,
                       verts, V, radius, table, 1, out);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

int surface_temp(nx_ctx *ctx, const double *verts, int64_t V, double radius, double tilt,
                 float *out, hipStream_t stream)
{
    if (!ctx) return set_error(NX_EINVAL, "null context");
    if (V < 0 || (V > 0 && (!verts || !out))) return set_error(NX_EINVAL, "bad vertices / output");
    if (V == 0) return NX_OK;
    NX_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_surface_temp, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, stream,
                       verts, V, radius, tilt, out);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

int yearly_insolation(nx_ctx *ctx, const double *verts, int64_t V, double radius,
                      double axial_tilt, float *out, void *ws, hipStream_t stream)
{
    int rc = climate_check(ctx, ws);
    if (rc) return rc;
    if (V < 0 || (V > 0 && (!verts || !out))) return set_error(NX_EINVAL, "bad vertices / output");
    NX_HIP(hipSetDevice(ctx->device));
    double tilts[kDays];
    for (int x = 0; x < kDays; ++x)  // climate.py:201
        tilts[x] = std::sin((double)x * 3.141592653589793 / 180) * axial_tilt;
    if ((rc = slices(ctx, radius, tilts, kDays, ws, stream))) return rc;
    if (V == 0) return NX_OK;
    hipLaunchKernelGGL(k_interpolate, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, stream,
                       verts, V, radius, (const float *)((char *)ws + climate_layout().tables),
                       kDays, out);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

}  // namespace nx
