// nx_api.cpp -- extern "C" entry points of libnixis_hip.so (include/nixis.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "nx_ctx.h"

namespace nx {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

// opensimplex.py:90-107.  over() is compiled by numba with the explicit
// signature int32(int32) (opensimplex.py:110-112), so each LCG step keeps only
// the low 32 bits, sign-extended: s <- int32(s * 6364136223846793005 +
// 1442695040888963407) with the int64 arithmetic wrapping.  The swap index
// r = (s + 31) mod (i + 1) is Python's floored modulo.  pgi is
// int((perm % 24.0) * 3), i.e. 3 * (perm mod 24).
int init_tables(int64_t seed, int32_t *perm, int32_t *pgi)
{
    if (!perm || !pgi) return set_error(NX_EINVAL, "perm/pgi must not be NULL");
    auto step = [](int64_t s) -> int64_t {
        const uint64_t t = (uint64_t)s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (int64_t)(int32_t)(uint32_t)t;
    };
    int32_t source[256];
    for (int i = 0; i < 256; ++i) source[i] = i;
    int64_t s = step(step(step(seed)));
    for (int i = 255; i >= 0; --i) {
        s = step(s);
        const int64_t m = i + 1;
        int64_t r = (s + 31) % m;
        if (r < 0) r += m;
        perm[i] = source[r];
        pgi[i] = 3 * (perm[i] % 24);
        source[r] = source[i];
    }
    return NX_OK;
}

char *stage_begin(nx_ctx *ctx, size_t bytes, int *slot)
{
    std::lock_guard<std::mutex> lk(ctx->stage_mu);
    const int i = (int)(ctx->stage_next++ % kStageSlots);
    StageSlot &s = ctx->stage[i];
    if (s.pending && hipEventSynchronize(s.ev) != hipSuccess) return nullptr;
    s.pending = false;
    if (!s.ev && hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess) return nullptr;
    if (s.cap < bytes) {
        if (s.host) (void)hipHostFree(s.host);
        s.host = nullptr;
        s.cap = 0;
        const size_t cap = std::max<size_t>(bytes, 64 << 10);
        if (hipHostMalloc(&s.host, cap, hipHostMallocDefault) != hipSuccess) return nullptr;
        s.cap = cap;
    }
    *slot = i;
    return (char *)s.host;
}

int stage_commit(nx_ctx *ctx, int slot, void *dst, size_t bytes, hipStream_t stream)
{
    StageSlot &s = ctx->stage[slot];
    NX_HIP(hipMemcpyAsync(dst, s.host, bytes, hipMemcpyHostToDevice, stream));
    NX_HIP(hipEventRecord(s.ev, stream));
    s.pending = true;
    return NX_OK;
}

}  // namespace nx

using namespace nx;

extern "C" {

const char *nx_last_error(void) { return g_err; }

int nx_abi_version(void) { return 1; }

int nx_ctx_create(int device, nx_ctx **out)
{
    if (!out) return set_error(NX_EINVAL, "out must not be NULL");
    *out = nullptr;
    int n = 0;
    NX_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return set_error(NX_EINVAL, "device %d outside 0..%d", device, n - 1);
    NX_HIP(hipSetDevice(device));
    nx_ctx *c = new (std::nothrow) nx_ctx();
    if (!c) return set_error(NX_ENOMEM, "context allocation failed");
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        c->num_cus = prop.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&c->capture_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&c->d_err, 256);
    if (e == hipSuccess) e = hipMalloc((void **)&c->d_keys, sizeof(unsigned long long) * 2 * kKeySlots);
    if (e == hipSuccess) e = hipMemset(c->d_err, 0, 256);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    // this device's address of the static chain scratch (device current)
    const int src = e == hipSuccess ? chain_scratch_init(c) : NX_OK;
    if (e != hipSuccess || src != NX_OK) {
        if (c->d_err) (void)hipFree(c->d_err);
        if (c->d_keys) (void)hipFree(c->d_keys);
        if (c->capture_stream) (void)hipStreamDestroy(c->capture_stream);
        delete c;
        if (src != NX_OK) return src;
        return set_error(NX_EHIP, "context device allocation failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return NX_OK;
}

int nx_ctx_destroy(nx_ctx *ctx)
{
    if (!ctx) return NX_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    ctx->clear_graphs();
    for (auto &s : ctx->stage) {
        if (s.ev) (void)hipEventDestroy(s.ev);
        if (s.host) (void)hipHostFree(s.host);
    }
    if (ctx->d_err) (void)hipFree(ctx->d_err);
    if (ctx->d_keys) (void)hipFree(ctx->d_keys);
    if (ctx->capture_stream) (void)hipStreamDestroy(ctx->capture_stream);
    delete ctx;
    return NX_OK;
}

int nx_ctx_set_option(nx_ctx *ctx, int option, int64_t value)
{
    if (!ctx) return set_error(NX_EINVAL, "null context");
    // Cached graphs stay valid: their keys hold the kernel variant and the
    // number of steps they cover, so no graph captured under other options
    // is replayed, and none that may still be running is destroyed here.
    std::lock_guard<std::mutex> lk(ctx->mu);
    switch (option) {
    case NX_OPT_GRAPH_CHUNK:
        if (value < 0 || value > 4096) return set_error(NX_EINVAL, "graph chunk %lld outside 0..4096", (long long)value);
        ctx->graph_chunk = (int)value;
        return NX_OK;
    case NX_OPT_MULTI_KERNEL:
        if (value < 0 || value > 4) return set_error(NX_EINVAL, "multi-step kernel variant %lld outside 0..4", (long long)value);
        ctx->multi_variant = (int)value;
        return NX_OK;
    case NX_OPT_PAIR_KERNEL:
        if (value < 0 || value > 6) return set_error(NX_EINVAL, "pair kernel variant %lld outside 0..6", (long long)value);
        ctx->pair_variant = (int)value;
        return NX_OK;
    case NX_OPT_CHAIN_PASSES:
        if (value != 2 && value != 5) return set_error(NX_EINVAL, "height chain passes %lld: 2 or 5", (long long)value);
        ctx->chain_passes = (int)value;
        return NX_OK;
    default:
        return set_error(NX_EINVAL, "unknown option %d", option);
    }
}

int nx_ctx_sync_check(nx_ctx *ctx, void *stream)
{
    if (!ctx) return set_error(NX_EINVAL, "null context");
    NX_HIP(hipSetDevice(ctx->device));
    NX_HIP(hipStreamSynchronize((hipStream_t)stream));
    int word = 0;
    std::lock_guard<std::mutex> lk(ctx->mu);
    NX_HIP(hipMemcpy(&word, ctx->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (word) {
        const int zero = 0;
        NX_HIP(hipMemcpy(ctx->d_err, &zero, sizeof(int), hipMemcpyHostToDevice));
    }
    if (word & NX_ERR_RANGE_BIT)
        return set_error(NX_ERANGE,
                         "noise coordinate outside the exact domain (|skewed coordinate| >= 2^30 "
                         "or not finite)");
    if (word & NX_ERR_DIV0_BIT) return set_error(NX_EDIV0, "rescale: input range is zero");
    if (word & NX_ERR_INDEX_BIT)
        return set_error(NX_EINVAL, "gray array: a neighbour index outside the colors array");
    if (word & NX_ERR_MESH_BIT)
        return set_error(NX_EINVAL, "mesh: a triangle index outside 0..V-1, a vertex in more than "
                                    "six triangles, or a neighbour ring that does not close");
    return NX_OK;
}

int nx_init_tables(int64_t seed, int32_t *perm, int32_t *pgi) { return init_tables(seed, perm, pgi); }

int nx_sample_octaves(nx_ctx *ctx, const double *verts, int64_t V, const int32_t *perm,
                      const int32_t *pgi, int n_octaves, const double *rough,
                      const double *strength, double radius, double *elev, int accumulate,
                      float *dbg_noise, int32_t *dbg_sel, void *stream)
{
    return sample_octaves(ctx, verts, V, perm, pgi, n_octaves, rough, strength, radius, elev,
                          accumulate, dbg_noise, dbg_sel, (hipStream_t)stream);
}

int nx_noise3d(nx_ctx *ctx, const double *x, const double *y, const double *z, int64_t n,
               const int32_t *perm, const int32_t *pgi, double *out, void *stream)
{
    return noise3d(ctx, x, y, z, n, perm, pgi, out, (hipStream_t)stream);
}

int nx_rescale(nx_ctx *ctx, const double *x, int64_t n, double lower, double upper, double *out,
               double *minmax, void *stream)
{
    return rescale(ctx, x, n, lower, upper, out, minmax, (hipStream_t)stream);
}

int nx_rescale_ex(nx_ctx *ctx, const double *x, int64_t n, double lower, double upper,
                  const nx_rescale_opts *opts, double *out, double *minmax, void *stream)
{
    return rescale_ex(ctx, x, n, lower, upper, opts, out, minmax, (hipStream_t)stream);
}

int nx_make_bool_elevation_mask(nx_ctx *ctx, const double *h, int64_t n, double level,
                                uint8_t *mask, void *stream)
{
    return make_mask(ctx, h, n, level, mask, (hipStream_t)stream);
}

size_t nx_power_rescale_workspace_bytes(int64_t n) { return power_rescale_workspace_bytes(n); }

int nx_power_rescale(nx_ctx *ctx, const double *x, const uint8_t *mask, int64_t n, int mode,
                     double power, double *out, void *ws, void *stream)
{
    return power_rescale(ctx, x, mask, n, mode, power, out, ws, (hipStream_t)stream);
}

size_t nx_height_chain_workspace_bytes(int64_t n) { return height_chain_workspace_bytes(n); }

size_t nx_climate_workspace_bytes(void) { return climate_workspace_bytes(); }

int nx_sample_insolation(nx_ctx *ctx, const double *verts, int64_t V, double radius,
                         double rotation, double rot_step, int nsteps, double tilt, float *arr,
                         void *ws, void *stream)
{
    return sample_insolation(ctx, verts, V, radius, rotation, rot_step, nsteps, tilt, arr, ws,
                             (hipStream_t)stream);
}

int nx_insolation_slice(nx_ctx *ctx, double radius, double tilt, float *out, void *ws,
                        void *stream)
{
    return insolation_slice(ctx, radius, tilt, out, ws, (hipStream_t)stream);
}

int nx_daily_insolation(nx_ctx *ctx, const double *verts, int64_t V, double radius, double tilt,
                        float *out, void *ws, void *stream)
{
    return daily_insolation(ctx, verts, V, radius, tilt, out, ws, (hipStream_t)stream);
}

int nx_interpolate_insolation(nx_ctx *ctx, const double *verts, int64_t V, const float *table,
                              double radius, float *out, void *stream)
{
    return interpolate_insolation(ctx, verts, V, table, radius, out, (hipStream_t)stream);
}

int nx_assign_surface_temp(nx_ctx *ctx, const double *verts, int64_t V, double radius,
                           double tilt, float *out, void *stream)
{
    return surface_temp(ctx, verts, V, radius, tilt, out, (hipStream_t)stream);
}

int nx_yearly_insolation(nx_ctx *ctx, const double *verts, int64_t V, double radius,
                         double axial_tilt, float *out, void *ws, void *stream)
{
    return yearly_insolation(ctx, verts, V, radius, axial_tilt, out, ws, (hipStream_t)stream);
}

size_t nx_knn_grid_bytes(int64_t V) { return knn_grid_bytes(V); }

int nx_knn_grid_build(nx_ctx *ctx, const double *points, int64_t V, void *grid, void *stream)
{
    return knn_grid_build(ctx, points, V, grid, (hipStream_t)stream);
}

int nx_knn_query(nx_ctx *ctx, const void *grid, int64_t V, const double *q, int64_t nq, int k,
                 double *dists, int64_t *nbrs, void *stream)
{
    return knn_query(ctx, grid, V, q, nq, k, dists, nbrs, (hipStream_t)stream);
}

size_t nx_pixel_workspace_bytes(int64_t width, int64_t height)
{
    return pixel_workspace_bytes(width, height);
}

int nx_knn_query_pixels(nx_ctx *ctx, const void *grid, int64_t V, int64_t width, int64_t height,
                        double radius, int k, double *dists, int64_t *nbrs, void *ws,
                        void *stream)
{
    return knn_query_pixels(ctx, grid, V, width, height, radius, k, dists, nbrs, ws,
                            (hipStream_t)stream);
}

int nx_make_ll_arr(nx_ctx *ctx, int64_t width, int64_t height, double radius, double *out,
                   void *ws, void *stream)
{
    return make_ll_arr(ctx, width, height, radius, out, ws, (hipStream_t)stream);
}

int nx_make_gray_array(nx_ctx *ctx, int64_t npix, int k, const double *dists,
                       const int64_t *nbrs, const double *colors, int64_t ncolors, int32_t *out,
                       void *stream)
{
    return make_gray_array(ctx, npix, k, dists, nbrs, colors, ncolors, out, (hipStream_t)stream);
}

int nx_terrain_heights(nx_ctx *ctx, const double *verts, int64_t V, const int32_t *perm,
                       const int32_t *pgi, int n_octaves, const double *rough,
                       const double *strength, double radius, double min_alt, double max_alt,
                       double ocean_percent, double *h, uint8_t *mask, double *ocean_level,
                       void *ws, void *stream)
{
    return terrain_heights(ctx, verts, V, perm, pgi, n_octaves, rough, strength, radius, min_alt,
                           max_alt, ocean_percent, h, mask, ocean_level, ws, (hipStream_t)stream);
}

int nx_terrain_heights_noalloc(nx_ctx *ctx, const double *verts, int64_t V, const int32_t *perm,
                               const int32_t *pgi, int n_octaves, const double *rough,
                               const double *strength, double radius, double min_alt,
                               double max_alt, double ocean_percent, double *h, uint8_t *mask,
                               double *ocean_level, void *stream)
{
    return terrain_heights_noalloc(ctx, verts, V, perm, pgi, n_octaves, rough, strength, radius,
                                   min_alt, max_alt, ocean_percent, h, mask, ocean_level,
                                   (hipStream_t)stream);
}

int nx_height_chain(nx_ctx *ctx, double *h, int64_t n, double min_alt, double max_alt,
                    double ocean_percent, uint8_t *mask, double *ocean_level, void *ws,
                    void *stream)
{
    return height_chain(ctx, h, n, min_alt, max_alt, ocean_percent, mask, ocean_level, ws,
                        (hipStream_t)stream);
}

size_t nx_erosion_state_bytes(int64_t V) { return erosion_state_bytes(V); }

int nx_erosion_load(nx_ctx *ctx, void *state, int64_t V, const double *h, const double *w,
                    const double *s, void *stream)
{
    return erosion_load(ctx, state, V, h, w, s, (hipStream_t)stream);
}

int nx_erosion_run(nx_ctx *ctx, void *state, int64_t V, const int32_t *col, const double *elen,
                   int iters, double rain, void *stream)
{
    return erosion_run(ctx, state, V, col, elen, nullptr, iters, rain, (hipStream_t)stream);
}

int nx_erosion_run_tiled(nx_ctx *ctx, void *state, int64_t V, const int32_t *col,
                         const double *elen, const nx_tile_plan *plan, int iters, double rain,
                         void *stream)
{
    return erosion_run(ctx, state, V, col, elen, plan, iters, rain, (hipStream_t)stream);
}

int nx_erosion_run_rule(nx_ctx *ctx, void *state, int64_t V, const int32_t *col,
                        const double *elen, const nx_tile_plan *plan, int iters, double rain,
                        int rule, void *stream)
{
    return erosion_run_rule(ctx, state, V, col, elen, plan, iters, rain, rule,
                            (hipStream_t)stream);
}

int nx_erosion1_run(nx_ctx *ctx, double *h, double *tmp, int64_t V, const int32_t *col, int iters,
                    void *stream)
{
    return erosion1_run(ctx, h, tmp, V, col, iters, (hipStream_t)stream);
}

int nx_icosphere_device(nx_ctx *ctx, int64_t d, double *verts, int32_t *tris, void *stream)
{
    return icosphere_dev(ctx, d, verts, tris, (hipStream_t)stream);
}

size_t nx_adjacency_workspace_bytes(int64_t V) { return adjacency_workspace_bytes(V); }

int nx_build_adjacency_device(nx_ctx *ctx, const int32_t *tris, int64_t T, int64_t V,
                              int32_t *adj, void *workspace, void *stream)
{
    return build_adjacency_dev(ctx, tris, T, V, adj, workspace, (hipStream_t)stream);
}

int nx_sort_adjacency_device(nx_ctx *ctx, const int32_t *adj_in, int32_t *adj_out, int64_t V,
                             void *stream)
{
    return sort_adjacency_dev(ctx, adj_in, adj_out, V, (hipStream_t)stream);
}

size_t nx_sfc_workspace_bytes(int64_t V) { return sfc_workspace_bytes(V); }

int nx_sfc_order_device(nx_ctx *ctx, const double *verts, int64_t V, int64_t *order,
                        void *workspace, void *stream)
{
    return sfc_order_dev(ctx, verts, V, order, workspace, (hipStream_t)stream);
}

size_t nx_ell_workspace_bytes(int64_t V) { return ell_workspace_bytes(V); }

int nx_ell_from_adjacency_device(nx_ctx *ctx, const int32_t *adj, int64_t V, const double *verts,
                                 const int64_t *order, int32_t *col, double *elen,
                                 void *workspace, void *stream)
{
    return ell_from_adjacency_dev(ctx, adj, V, verts, order, col, elen, workspace,
                                  (hipStream_t)stream);
}

int nx_tile_plan_build(const int32_t *col, const double *elen, int64_t V, void *out, int64_t cap,
                       nx_tile_plan *plan)
{
    return tile_plan_build(col, elen, V, 2, out, cap, plan);
}

int nx_tile_plan_build_steps(const int32_t *col, const double *elen, int64_t V, int steps,
                             void *out, int64_t cap, nx_tile_plan *plan)
{
    return tile_plan_build(col, elen, V, steps, out, cap, plan);
}

int nx_tile_plan_build_chunk(const int32_t *col, const double *elen, int64_t V, int steps,
                             int chunk, void *out, int64_t cap, nx_tile_plan *plan)
{
    return tile_plan_build(col, elen, V, steps, out, cap, plan, chunk);
}

int nx_erosion_store(nx_ctx *ctx, const void *state, int64_t V, double *h, double *w, double *s,
                     void *stream)
{
    return erosion_store(ctx, state, V, h, w, s, (hipStream_t)stream);
}

int nx_erosion_mass(nx_ctx *ctx, const void *state, int64_t V, double *out, void *stream)
{
    return erosion_mass(ctx, state, V, out, (hipStream_t)stream);
}

size_t nx_batch_workspace_bytes(int n) { return batch_workspace_bytes(n); }

int nx_batch_noise(nx_ctx *ctx, nx_planet *planets, int n, void *workspace, int tile,
                   int blocks_per_cu, void *stream)
{
    return batch_noise(ctx, planets, n, workspace, tile, blocks_per_cu, (hipStream_t)stream);
}

int nx_batch_erode(nx_ctx *ctx, nx_planet *planets, int n, int max_steps, void *workspace,
                   int use_graph, void *stream)
{
    return batch_erode(ctx, planets, n, max_steps, workspace, use_graph, (hipStream_t)stream);
}

int nx_batch_store(nx_ctx *ctx, const nx_planet *planets, int n, void *workspace, void *stream)
{
    return batch_store(ctx, planets, n, workspace, (hipStream_t)stream);
}

int nx_icosphere_counts(int64_t d, int64_t *V, int64_t *T)
{
    const int rc = icosphere_counts(d, V, T);
    return rc ? set_error(rc, "divisions %lld outside 1..100000", (long long)d) : NX_OK;
}

int nx_icosphere(int64_t d, double *verts, int32_t *tris)
{
    const int rc = icosphere(d, verts, tris);
    return rc ? set_error(rc, "divisions %lld outside 1..14650 (V must fit int32)", (long long)d)
              : NX_OK;
}

int nx_build_adjacency(const int32_t *tris, int64_t T, int64_t V, int32_t *adj)
{
    const int rc = build_adjacency(tris, T, V, adj);
    return rc ? set_error(rc, "build_adjacency: bad triangle list (index or degree > 6)") : NX_OK;
}

int nx_sort_adjacency(int32_t *adj, int64_t V)
{
    const int rc = sort_adjacency(adj, V);
    return rc ? set_error(rc, "sort_adjacency: a row is not a closed ring") : NX_OK;
}

int nx_sfc_order(const double *verts, int64_t V, int64_t *order)
{
    const int rc = sfc_order(verts, V, order);
    return rc ? set_error(rc, "sfc_order: bad arguments") : NX_OK;
}

int nx_ell_from_adjacency(const int32_t *adj, int64_t V, const double *verts, const int64_t *order,
                          int32_t *col, double *elen)
{
    const int rc = ell_from_adjacency(adj, V, verts, order, col, elen);
    return rc ? set_error(rc, "ell_from_adjacency: bad table or order") : NX_OK;
}

}  // extern "C"
