// nx_ctx.h -- the context object behind nx_ctx* and HIP error plumbing.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <mutex>
#include <tuple>

#include "nx_internal.h"

// Bits of the per-context sticky error word written by kernels.
#define NX_ERR_RANGE_BIT 1
#define NX_ERR_DIV0_BIT 2
#define NX_ERR_MESH_BIT 4  // device mesh builders: bad index, degree > 6 or broken ring
#define NX_ERR_INDEX_BIT 8  // gray sampling: a neighbour index outside the colors

constexpr int kKeySlots = 256;
constexpr size_t kMaxGraphs = 64;

// Identity of a captured chunk of erosion steps.
struct GraphKey {
    const void *state;
    int64_t V;
    const void *col;
    const void *elen;
    int n;
    double rain;
    bool last;
    uint64_t tag;  // hash of what the captured launches take by value (plan, schedule)
    bool operator<(const GraphKey &o) const
    {
        return std::tie(state, V, col, elen, n, rain, last, tag) <
               std::tie(o.state, o.V, o.col, o.elen, o.n, o.rain, o.last, o.tag);
    }
};

// Pinned host staging buffer for descriptor uploads; `ev` marks the end of
// the last copy out of it.
struct StageSlot {
    void *host = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
};
constexpr int kStageSlots = 8;

struct nx_ctx {
    int device = 0;
    int num_cus = 256;
    int *d_err = nullptr;                  // device word, bits NX_ERR_*
    unsigned long long *d_keys = nullptr;  // kKeySlots x 2 reduction keys
    std::atomic<unsigned> key_slot{0};
    hipStream_t capture_stream = nullptr;  // default priority, graph capture only
    int graph_chunk = 32;                  // erosion steps per captured graph; 0 = eager
    int pair_variant = 0;                  // two-step kernel variant (0 = default)
    int multi_variant = 0;                 // four-step kernel variant (0 = default)
    int chain_passes = 2;                  // height chain: 2 (default) or the five-pass form
    void *chain_scratch = nullptr;         // this device's address of the static chain scratch
    std::map<GraphKey, hipGraphExec_t> graphs;
    std::mutex mu;
    StageSlot stage[kStageSlots];
    unsigned stage_next = 0;
    std::mutex stage_mu;

    void clear_graphs()
    {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
    // A full cache is emptied once the device has finished every launch, so
    // no graph that may still be running is destroyed.
    void make_room_for_graph()
    {
        if (graphs.size() < kMaxGraphs) return;
        (void)hipSetDevice(device);
        (void)hipDeviceSynchronize();
        clear_graphs();
    }
};

#define NX_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::nx::set_error(NX_EHIP, "%s failed: %s (%s:%d)", #call,               \
                                   hipGetErrorString(e_), __FILE__, __LINE__);            \
    } while (0)

namespace nx {

int sample_octaves(nx_ctx *ctx, const double *verts, int64_t V, const int32_t *perm,
                   const int32_t *pgi, int n_oct, const double *rough, const double *strength,
                   double radius, double *elev, int accumulate, float *dbg_noise,
                   int32_t *dbg_sel, hipStream_t stream);
int noise3d(nx_ctx *ctx, const double *x, const double *y, const double *z, int64_t n,
            const int32_t *perm, const int32_t *pgi, double *out, hipStream_t stream);
int rescale(nx_ctx *ctx, const double *x, int64_t n, double lower, double upper, double *out,
            double *minmax, hipStream_t stream);
int rescale_ex(nx_ctx *ctx, const double *x, int64_t n, double lower, double upper,
               const nx_rescale_opts *o, double *out, double *minmax, hipStream_t stream);
int make_mask(nx_ctx *ctx, const double *h, int64_t n, double level, uint8_t *mask,
              hipStream_t stream);
size_t power_rescale_workspace_bytes(int64_t n);
int power_rescale(nx_ctx *ctx, const double *x, const uint8_t *mask, int64_t n, int mode,
                  double power, double *out, void *ws, hipStream_t stream);
int sample_octaves_keys(nx_ctx *ctx, const double *verts, int64_t V, const int32_t *perm,
                        const int32_t *pgi, int n_oct, const double *rough,
                        const double *strength, double radius, double *elev, int accumulate,
                        float *dbg_noise, int32_t *dbg_sel, unsigned long long *keys,
                        hipStream_t stream);
int terrain_heights(nx_ctx *ctx, const double *verts, int64_t V, const int32_t *perm,
                    const int32_t *pgi, int n_oct, const double *rough, const double *strength,
                    double radius, double min_alt, double max_alt, double ocean_percent,
                    double *h, uint8_t *mask, double *level, void *ws, hipStream_t stream);
size_t height_chain_workspace_bytes(int64_t n);
int chain_scratch_init(nx_ctx *ctx);
int terrain_heights_noalloc(nx_ctx *ctx, const double *verts, int64_t V, const int32_t *perm,
                            const int32_t *pgi, int n_oct, const double *rough,
                            const double *strength, double radius, double min_alt,
                            double max_alt, double ocean_percent, double *h, uint8_t *mask,
                            double *level, hipStream_t stream);
size_t climate_workspace_bytes();
int sample_insolation(nx_ctx *ctx, const double *verts, int64_t V, double radius, double rot0,
                      double rot_step, int nsteps, double tilt, float *arr, void *ws,
                      hipStream_t stream);
int insolation_slice(nx_ctx *ctx, double radius, double tilt, float *out, void *ws,
                     hipStream_t stream);
int daily_insolation(nx_ctx *ctx, const double *verts, int64_t V, double radius, double tilt,
                     float *out, void *ws, hipStream_t stream);
int surface_temp(nx_ctx *ctx, const double *verts, int64_t V, double radius, double tilt,
                 float *out, hipStream_t stream);
int interpolate_insolation(nx_ctx *ctx, const double *verts, int64_t V, const float *table,
                           double radius, float *out, hipStream_t stream);
int yearly_insolation(nx_ctx *ctx, const double *verts, int64_t V, double radius,
                      double axial_tilt, float *out, void *ws, hipStream_t stream);
size_t knn_grid_bytes(int64_t V);
int knn_grid_build(nx_ctx *ctx, const double *pts, int64_t V, void *grid, hipStream_t stream);
int knn_query(nx_ctx *ctx, const void *grid, int64_t V, const double *q, int64_t nq, int k,
              double *dists, int64_t *nbrs, hipStream_t stream);
size_t pixel_workspace_bytes(int64_t width, int64_t height);
int knn_query_pixels(nx_ctx *ctx, const void *grid, int64_t V, int64_t width, int64_t height,
                     double radius, int k, double *dists, int64_t *nbrs, void *ws,
                     hipStream_t stream);
int make_ll_arr(nx_ctx *ctx, int64_t width, int64_t height, double radius, double *out, void *ws,
                hipStream_t stream);
int make_gray_array(nx_ctx *ctx, int64_t npix, int k, const double *dists, const int64_t *nbrs,
                    const double *colors, int64_t ncolors, int32_t *out, hipStream_t stream);
int height_chain(nx_ctx *ctx, double *h, int64_t n, double min_alt, double max_alt,
                 double ocean_percent, uint8_t *mask, double *level, void *ws, hipStream_t stream);
size_t erosion_state_bytes(int64_t V);
int erosion_load(nx_ctx *ctx, void *state, int64_t V, const double *h, const double *w,
                 const double *s, hipStream_t stream);
int erosion_run(nx_ctx *ctx, void *state, int64_t V, const int32_t *col, const double *elen,
                const nx_tile_plan *plan, int iters, double rain, hipStream_t stream);
int erosion_run_rule(nx_ctx *ctx, void *state, int64_t V, const int32_t *col, const double *elen,
                     const nx_tile_plan *plan, int iters, double rain, int rule,
                     hipStream_t stream);
int erosion1_run(nx_ctx *ctx, double *h, double *tmp, int64_t V, const int32_t *col, int iters,
                 hipStream_t stream);
int erosion_store(nx_ctx *ctx, const void *state, int64_t V, double *h, double *w, double *s,
                  hipStream_t stream);
int erosion_mass(nx_ctx *ctx, const void *state, int64_t V, double *out, hipStream_t stream);
// Host-to-device upload through the context's ring of pinned staging
// buffers, so that it is asynchronous (a copy from pageable memory is not).
// stage_begin returns a pinned buffer of at least `bytes` for slot *slot,
// waiting first for that slot's previous copy; stage_commit copies it to
// `dst` on `stream` and records the slot's event.
char *stage_begin(nx_ctx *ctx, size_t bytes, int *slot);
int stage_commit(nx_ctx *ctx, int slot, void *dst, size_t bytes, hipStream_t stream);
int icosphere_dev(nx_ctx *ctx, int64_t d, double *verts, int32_t *tris, hipStream_t stream);
size_t adjacency_workspace_bytes(int64_t V);
int build_adjacency_dev(nx_ctx *ctx, const int32_t *tris, int64_t T, int64_t V, int32_t *adj,
                        void *workspace, hipStream_t stream);
int sort_adjacency_dev(nx_ctx *ctx, const int32_t *adj_in, int32_t *adj_out, int64_t V,
                       hipStream_t stream);
size_t sfc_workspace_bytes(int64_t V);
int sfc_order_dev(nx_ctx *ctx, const double *verts, int64_t V, int64_t *order, void *ws,
                  hipStream_t stream);
size_t ell_workspace_bytes(int64_t V);
int ell_from_adjacency_dev(nx_ctx *ctx, const int32_t *adj, int64_t V, const double *verts,
                           const int64_t *order, int32_t *col, double *elen, void *ws,
                           hipStream_t stream);
size_t batch_workspace_bytes(int n);
int batch_noise(nx_ctx *ctx, nx_planet *planets, int n, void *ws, int tile, int blocks_per_cu,
                hipStream_t stream);
int batch_erode(nx_ctx *ctx, nx_planet *planets, int n, int max_steps, void *ws, int use_graph,
                hipStream_t stream);
int batch_store(nx_ctx *ctx, const nx_planet *planets, int n, void *ws, hipStream_t stream);

}  // namespace nx
