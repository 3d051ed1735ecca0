/*
 * nixis.h -- C ABI of libnixis_hip.so, the MI355X (gfx950) implementation of
 * the Nixis terrain hot path: fractal OpenSimplex noise at every icosphere
 * vertex and the height/water/sediment erosion iteration.
 *
 * Conventions
 *   - Every entry point returns an int status: NX_OK (0) or a negative
 *     NX_E* code; nx_last_error() then returns a message (thread-local).
 *   - Arrays marked "device" are caller-owned device pointers (for example
 *     torch tensors' data_ptr()).  Arrays marked "host" are ordinary host
 *     memory.  Nothing here takes ownership of caller memory.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Device entry points are asynchronous on that stream.  Range errors the
 *     kernels detect (NX_ERANGE, NX_EDIV0) are sticky per context and are
 *     returned by nx_ctx_sync_check().
 *   - Library-owned state (sticky error word, graphs, autotune results) hangs
 *     off an opaque nx_ctx, one per device.
 *
 * The reference (MightyBOBcnc/nixis) is pure Python + numba and has no FFI;
 * each entry point names the reference function (file:line) whose behaviour
 * it provides.  The Python modules in nixis_amd/ bind these through ctypes
 * under the reference's own function names and signatures.
 */
#ifndef NIXIS_H
#define NIXIS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NX_OK 0
#define NX_EINVAL (-1)    /* bad argument (size, null pointer, range)      */
#define NX_ENOMEM (-2)    /* allocation failed                              */
#define NX_EHIP (-3)      /* a HIP runtime call failed                      */
#define NX_ERANGE (-4)    /* input outside the kernel's exact domain        */
#define NX_EDIV0 (-5)     /* zero input range in rescale                    */
#define NX_EBUSY (-6)     /* serving queue full / HBM budget exhausted      */

#define NX_MAX_OCTAVES 32 /* octaves per sample_octaves call               */

typedef struct nx_ctx nx_ctx;

/* ---- library ----------------------------------------------------------- */

/* Message for the last failing call on this host thread ("" if none). */
const char *nx_last_error(void);
/* ABI version, bumped on any signature change. */
int nx_abi_version(void);

/* Create / destroy a context bound to HIP device `device`. */
int nx_ctx_create(int device, nx_ctx **out);
int nx_ctx_destroy(nx_ctx *ctx);
/* Context options (nx_ctx_set_option). */
#define NX_OPT_GRAPH_CHUNK 1 /* erosion steps per captured hipGraph; 0 = eager
                                launches (default 32)                        */
#define NX_OPT_PAIR_KERNEL 2 /* two-step erosion kernel variant, 0..6 (0 =
                                default); a tuning knob, results identical   */
#define NX_OPT_MULTI_KERNEL 3 /* four-step erosion kernel variant, 0..4 (0 =
                               * default: first-step data kept in registers,
                               * two LDS buffers; 1..4 = slot tables in LDS,
                               * 1024 or 512 threads)                       */
#define NX_OPT_CHAIN_PASSES 4 /* height chain form: 2 (default: one summary
                               * pass, one writing pass) or 5 (the pass per
                               * step form); results identical               */
int nx_ctx_set_option(nx_ctx *ctx, int option, int64_t value);
/* Synchronise `stream`, then return and clear the context's sticky kernel
 * error (NX_ERANGE: a noise coordinate outside |x| < 2^30 after the octave
 * multiplier, or not finite; NX_EDIV0: rescale of a constant array). */
int nx_ctx_sync_check(nx_ctx *ctx, void *stream);

/* ---- permutation tables (host) ----------------------------------------- */

/* opensimplex.py:90-107 (init) with over() at opensimplex.py:110-112:
 * perm/pgi are host int32[256]. */
int nx_init_tables(int64_t seed, int32_t *perm, int32_t *pgi);

/* ---- noise (device) ---------------------------------------------------- */

/* terrain.py:32-59 (sample_octaves) over terrain.py:12-28 (sample_noise) and
 * opensimplex.py:266-759 (noise3d), all octaves fused in one launch.
 *   verts     device f64 [V,3]
 *   perm,pgi  host int32[256] from nx_init_tables
 *   rough     host f64 [n_octaves]: coordinate multiplier of each octave
 *             (the reference's n_freq / world_radius)
 *   strength  host f64 [n_octaves]: n_amp / world_radius
 *   radius    world_radius
 *   elev      device f64 [V]; with accumulate != 0 the octaves are added to
 *             its contents (the reference's elevations argument), otherwise
 *             it is overwritten.
 *   dbg_noise device f32 [n_octaves,V] or NULL: per-octave noise3d value
 *   dbg_sel   device int32 [n_octaves,V,5] or NULL: per-octave lattice
 *             selection (xsb, ysb, zsb, region/extra-vertex code, checksum
 *             of the contributing gradient indices)
 * Lattice skew, floor, region and hash selection run in FP64 / integers and
 * are bit-exact with the reference; the gradient and falloff arithmetic runs
 * in FP32 (|error| <= 1e-5 of the FP64 noise value). */
int nx_sample_octaves(nx_ctx *ctx, const double *verts, int64_t V,
                      const int32_t *perm, const int32_t *pgi, int n_octaves,
                      const double *rough, const double *strength,
                      double radius, double *elev, int accumulate,
                      float *dbg_noise, int32_t *dbg_sel, void *stream);

/* opensimplex.py:257-263 (noisearr3d) / 266-759 (noise3d) at n points given
 * as device f64 arrays x, y, z; out is device f64 [n]. */
int nx_noise3d(nx_ctx *ctx, const double *x, const double *y, const double *z,
               int64_t n, const int32_t *perm, const int32_t *pgi,
               double *out, void *stream);

/* ---- rescale (device) -------------------------------------------------- */

/* util.py:110-140 (rescale, standard mode): out = ((x - min) / (max - min))
 * * (upper - lower) + lower with min/max reduced on the device.  x and out
 * are device f64 [n] and may alias.  minmax (device f64[2], nullable)
 * receives (min, max). */
int nx_rescale(nx_ctx *ctx, const double *x, int64_t n, double lower,
               double upper, double *out, double *minmax, void *stream);

/* util.py:110-175 (rescale with every option).  opts == NULL is the
 * standard form.  has_mid with mode 0 rescales [min, mid] onto
 * [lower, mid] and (mid, max] onto (mid, upper]; mode 1 ('lower') maps only
 * the values <= mid, mode 2 ('upper') only the values >= mid, copying the
 * rest; u_min / u_max widen the input range, never narrow it.  A mode
 * without mid is NX_EINVAL (the reference prints an error and returns x).
 * A zero range that a value divides by is NX_EDIV0 at the next sync.
 * minmax (nullable) receives the input's (min, max) before u_min/u_max. */
typedef struct nx_rescale_opts {
    double mid, u_min, u_max;
    int has_mid;
    int mode;                /* 0 none, 1 'lower', 2 'upper' */
    int has_umin, has_umax;
} nx_rescale_opts;

int nx_rescale_ex(nx_ctx *ctx, const double *x, int64_t n, double lower,
                  double upper, const nx_rescale_opts *opts, double *out,
                  double *minmax, void *stream);

/* terrain.py:61-72: mask[i] = h[i] <= level (u8 0/1). */
int nx_make_bool_elevation_mask(nx_ctx *ctx, const double *h, int64_t n,
                                double level, uint8_t *mask, void *stream);

/* util.py:178-254 (power_rescale).  mode 1 rescales the values whose mask is
 * set, mode 0 the others, through [0, 1] ** power and back, over the range
 * the reference's sequential loop finds: lower = min of the selected values
 * (starting from max(x)); upper = max of the selected values that were not
 * a new running minimum when the loop reached them (starting from min(x)),
 * found from in-order chunk summaries on the device.  x and out may alias.
 * ws: device workspace of nx_power_rescale_workspace_bytes(n) bytes,
 * 256-byte aligned. */
size_t nx_power_rescale_workspace_bytes(int64_t n);
int nx_power_rescale(nx_ctx *ctx, const double *x, const uint8_t *mask,
                     int64_t n, int mode, double power, double *out, void *ws,
                     void *stream);

/* nixis.py:330-364: the height chain after noise, in place on h:
 *   h = rescale(h, min_alt, max_alt)
 *   ocean_level = find_percent_val(min(h), max(h), ocean_percent)  (util.py:556-566)
 *   mask = h <= ocean_level
 *   h = power_rescale(h, mask, 1, 0.5); h = power_rescale(h, mask, 0, 2.0)
 *   h -= ocean_level; h = rescale(h, min_alt, max_alt, mid=0)
 * ocean_level (device f64[1]) and mask (device u8 [n]) are outputs; nothing
 * returns to the host.  ws: nx_height_chain_workspace_bytes(n), 256-aligned. */
size_t nx_height_chain_workspace_bytes(int64_t n);
int nx_height_chain(nx_ctx *ctx, double *h, int64_t n, double min_alt,
                    double max_alt, double ocean_percent, uint8_t *mask,
                    double *ocean_level, void *ws, void *stream);

/* terrain.py:32-59 then nixis.py:330-364 in one call: the noise heights at
 * the vertices (starting from zero, as nixis.py:330 passes no elevations)
 * with their min/max reduced inside the noise kernel, then the height chain
 * on them.  Arguments as nx_sample_octaves and nx_height_chain; h (device f64
 * [V]) receives the final heights.  Results equal nx_sample_octaves followed
 * by nx_height_chain.  ws: nx_height_chain_workspace_bytes(V). */
int nx_terrain_heights(nx_ctx *ctx, const double *verts, int64_t V,
                       const int32_t *perm, const int32_t *pgi, int n_octaves,
                       const double *rough, const double *strength,
                       double radius, double min_alt, double max_alt,
                       double ocean_percent, double *h, uint8_t *mask,
                       double *ocean_level, void *ws, void *stream);

/* nx_terrain_heights without a workspace argument.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * It calls no hipMalloc, hipFree or allocating library routine; the chain's
 * scratch (its 256-byte head and at most 2048 pairs of block summaries,
 * 196,864 bytes, enough for any V) is a static device array, one per GPU,
 * held until the process ends, whose address the context takes when it is
 * created.  Calls that overlap on one GPU share that array and race (also
 * from different contexts); calls in stream order are independent.
 * Arguments and results as nx_terrain_heights. */
int nx_terrain_heights_noalloc(nx_ctx *ctx, const double *verts, int64_t V,
                               const int32_t *perm, const int32_t *pgi,
                               int n_octaves, const double *rough,
                               const double *strength, double radius,
                               double min_alt, double max_alt,
                               double ocean_percent, double *h, uint8_t *mask,
                               double *ocean_level, void *stream);

/* ---- equirectangular export (device) ----------------------------------- */

/* util.py:275-287 (build_KDTree) and nixis.py:282 (KDTree.query(ll, k=3)):
 * the k nearest points of each query point, nearest first, as scipy's
 * KDTree returns them: dists = sqrt of the squared differences summed in
 * x, y, z order (f64), nbrs = point indices (int64); equal distances are
 * ordered by index.  Where fewer than k points exist the rest is +inf / V.
 * The structure is a uniform grid of cells in a device workspace of
 * nx_knn_grid_bytes(V) bytes (256-byte aligned): points sorted by cell and
 * a hash table of occupied cells.  1 <= k <= 8, V < 2^31 - 1. */
size_t nx_knn_grid_bytes(int64_t V);
int nx_knn_grid_build(nx_ctx *ctx, const double *points, int64_t V,
                      void *grid, void *stream);
int nx_knn_query(nx_ctx *ctx, const void *grid, int64_t V, const double *q,
                 int64_t nq, int k, double *dists, int64_t *nbrs, void *stream);

/* util.py:291-310 (make_ll_arr) and the query of its pixels in one pass:
 * pixel (h, w) of a width x height map sits at latitude
 * max(90 - h * 180 / (height - 1), -90) and longitude
 * max(-180 + w * 360 / width, -180), at (r cos lat cos lon,
 * r cos lat sin lon, r sin lat) (util.py:80-90); the cosines and sines of
 * each row and column are computed on the host with the C library's cos /
 * sin, so the coordinates equal the reference's.  dists / nbrs are
 * [height][width][k].  ws: nx_pixel_workspace_bytes(width, height). */
size_t nx_pixel_workspace_bytes(int64_t width, int64_t height);
int nx_knn_query_pixels(nx_ctx *ctx, const void *grid, int64_t V,
                        int64_t width, int64_t height, double radius, int k,
                        double *dists, int64_t *nbrs, void *ws, void *stream);
/* util.py:291-310: out is device f64 [height][width][3]. */
int nx_make_ll_arr(nx_ctx *ctx, int64_t width, int64_t height, double radius,
                   double *out, void *ws, void *stream);

/* util.py:343-366 (make_gray_array, and make_rgb_array's value,
 * util.py:313-340): for each of npix pixels with k neighbours,
 * sd = sum(d), w_i = 1 / ((d_i + 0.00001) / sd), t = sum(w),
 * out = (int32)(sum colors[n_i] * (w_i / t)), sums in neighbour order,
 * truncated toward zero.  colors: device f64 [ncolors]; a neighbour index
 * outside it is NX_EINVAL at the next sync. */
int nx_make_gray_array(nx_ctx *ctx, int64_t npix, int k, const double *dists,
                       const int64_t *nbrs, const double *colors,
                       int64_t ncolors, int32_t *out, void *stream);

/* ---- insolation (device) ------------------------------------------------ */

/* Insolation at the top of the atmosphere (climate.py:416-590).  Vertex
 * coordinates are device f64 [V][3]; results are device float32, summed as
 * numba adds a double to a float32 element (double sum, rounded to float).
 * The rotation, tilt and seasonal sines and cosines come from the host C
 * library; per-vertex asin / atan2 / cos run on the device, so values agree
 * with the reference's to a few float32 ulp per summed step.
 * ws: nx_climate_workspace_bytes() of 256-byte aligned device memory. */
size_t nx_climate_workspace_bytes(void);
/* climate.py:416-446 (sample_insolation) for nsteps (<= 360) rotations from
 * `rotation`, adding rot_step after each (climate.py:468-469), into arr;
 * nsteps = 1 is calc_instant_insolation (climate.py:489-498), 360 steps of
 * 1 degree from -180 brute_daily_insolation (climate.py:449-486). */
int nx_sample_insolation(nx_ctx *ctx, const double *verts, int64_t V,
                         double radius, double rotation, double rot_step,
                         int nsteps, double tilt, float *arr, void *ws,
                         void *stream);
/* climate.py:503-533 (calc_insolation_slice): out is device float32 [181],
 * index = latitude, negative latitudes at 181 + lat. */
int nx_insolation_slice(nx_ctx *ctx, double radius, double tilt, float *out,
                        void *ws, void *stream);
/* climate.py:547-573 (interpolate_insolation) with a device table [181]. */
int nx_interpolate_insolation(nx_ctx *ctx, const double *verts, int64_t V,
                              const float *table, double radius, float *out,
                              void *stream);
/* climate.py:351-412 (assign_surface_temp): out float32 [V] =
 * max(cos(|lat - tilt| * pi / 180), 0), the branch every vertex takes with
 * the reference's altitude weight of 0.  The caller raises the reference's
 * ZeroDivisionError for constant altitudes (its rescale, climate.py:374). */
int nx_assign_surface_temp(nx_ctx *ctx, const double *verts, int64_t V,
                           double radius, double tilt, float *out, void *stream);
/* climate.py:536-544 (calc_daily_insolation): slice + interpolation. */
int nx_daily_insolation(nx_ctx *ctx, const double *verts, int64_t V,
                        double radius, double tilt, float *out, void *ws,
                        void *stream);
/* climate.py:576-590 (calc_yearly_insolation): 360 days, seasonal tilt
 * sin(day * pi / 180) * axial_tilt (climate.py:193-201), daily values
 * summed in float32 in day order. */
int nx_yearly_insolation(nx_ctx *ctx, const double *verts, int64_t V,
                         double radius, double axial_tilt, float *out,
                         void *ws, void *stream);

/* ---- erosion (device) -------------------------------------------------- */

/* Fields live in a caller-allocated device workspace of
 * nx_erosion_state_bytes(V) bytes: height, water and sediment in FP64,
 * double-buffered.  Every step follows erosion.py:209-267 operation by
 * operation, so results equal the reference's FP64 results. */
size_t nx_erosion_state_bytes(int64_t V);

/* Load FP64 fields (device f64 [V]; w and s may be NULL for zeros). */
int nx_erosion_load(nx_ctx *ctx, void *state, int64_t V, const double *h,
                    const double *w, const double *s, void *stream);

/* erosion.py:171-192 (erode_terrain3) over erosion.py:197-279
 * (erosion_iteration3): `iters` Jacobi steps; before each step `rain` is
 * added to every vertex's water (the reference's water += 0.3/320; pass 0
 * for a bare erosion_iteration3).  Multi-step runs replay captured
 * hipGraphs of NX_OPT_GRAPH_CHUNK steps, cached per context.
 *   col   device int32 [6,V]  neighbour of each slot (self for a pad slot)
 *   elen  device f64 [6,V]    edge length of each slot (0 for a pad) */
int nx_erosion_run(nx_ctx *ctx, void *state, int64_t V, const int32_t *col,
                   const double *elen, int iters, double rain, void *stream);

/* Tile plan for running erosion steps several per launch (2, or 4 with
 * nx_tile_plan_build_steps / _chunk).  The vertices (in device order) are cut into
 * contiguous tiles of at most max_tile vertices whose successive layers of
 * outside neighbours (ring 1, ring 2, ...) bring the total to at most
 * max_load; each tile carries its rings' global indices and, for the tile
 * and all rings but the last, the neighbour slots renumbered into that local
 * set plus their edge lengths, each undirected edge's length stored once
 * (a slot names the column, and whether its own vertex or the neighbour
 * holds it).  The plan is a byte image: nx_tile_plan_build
 * fills it on the host, the caller copies it to device memory (256-byte
 * aligned) and stores that address in `blob`.  The same results as
 * nx_erosion_run (erosion.py:197-279 per step), with about half the HBM
 * traffic per step. */
typedef struct nx_tile_plan {
    void *blob;          /* device copy of the plan image (set by the caller) */
    int64_t bytes;       /* size of the plan image */
    int64_t V;
    int64_t ntiles;
    int32_t max_tile;    /* largest tile (owned vertices): <= 768 (2 steps), 1024 (4) */
    int32_t max_load;    /* largest tile + rings, <= 2048 (2 steps), 3072 (4) */
    int64_t off_start, off_cs, off_load, off_rbase, off_ring, off_cbase,
            off_lcol, off_lel;   /* byte offsets of the arrays in the image */
    int32_t steps;       /* steps per launch the plan serves: 2 or 4 */
    int32_t max_step1;   /* largest first-step set: tile + rings 1..steps-1 */
    int64_t off_pref;    /* int32 [ntiles][steps + 1]: tile, tile + ring 1, ... */
    int64_t off_ebase;   /* int64 [ntiles + 1]: each tile's edge lengths in lel */
} nx_tile_plan;

/* Build the plan from the host copies of col [6,V] and elen [6,V] (as
 * written by nx_ell_from_adjacency).  With out == NULL or cap < the size,
 * only fills `plan` (bytes, counts, offsets) and returns NX_OK; call again
 * with a buffer of plan->bytes.  Returns NX_ERANGE when some tile cannot be
 * cut small enough for its rings to fit (a graph without locality): no plan,
 * use nx_erosion_run. */
int nx_tile_plan_build(const int32_t *col, const double *elen, int64_t V,
                       void *out, int64_t cap, nx_tile_plan *plan);

/* nx_tile_plan_build for `steps` (2 or 4) steps per launch: tiles carry
 * `steps` rings; cs counts the tile and rings 1..steps-1 (the first step's
 * vertices), load all of them.  Two steps: tile + ring 1 at most 1024 (one
 * thread each in the two-step kernel) and tile + rings at most 2048.  Four
 * steps: tile + rings 1..3 at most 2048 (two per thread of the four-step
 * kernel) and tile + rings at most 2688.  A 4-step plan runs four steps per
 * launch in nx_erosion_run_tiled / _rule and nx_batch_erode (the neighbour
 * data read once per launch serves four steps: about 40% fewer bytes per
 * step than 2-step plans at d=1000, and 30% less time). */
int nx_tile_plan_build_steps(const int32_t *col, const double *elen, int64_t V,
                             int steps, void *out, int64_t cap,
                             nx_tile_plan *plan);

/* nx_tile_plan_build_steps with the starting tile size `chunk` (kTileMin =
 * 32 .. 1024 vertices for 2 steps, .. 2048 for 4; 0 for the default, 768 for
 * 2 steps and 1152 for 4).
 * Tiles are halved from there until their rings fit, so the chunk trades
 * ring overhead against the number of blocks per launch; the erosion
 * autotuner picks it per mesh size.  The results of erosion do not depend on
 * it. */
int nx_tile_plan_build_chunk(const int32_t *col, const double *elen, int64_t V,
                             int steps, int chunk, void *out, int64_t cap,
                             nx_tile_plan *plan);

/* nx_erosion_run with a tile plan (NULL plan: exactly nx_erosion_run).
 * Steps go plan->steps per launch; with a 4-step plan a remainder of 2 or 3
 * steps takes one launch, and a last single step runs alone. */
int nx_erosion_run_tiled(nx_ctx *ctx, void *state, int64_t V,
                         const int32_t *col, const double *elen,
                         const nx_tile_plan *plan, int iters, double rain,
                         void *stream);

/* erosion.py:301-381 (erosion_iteration4): rule 4 runs the same step as
 * rule 3 (erosion_iteration3) with every neighbour contribution and every
 * update clamped by max(x, 0) as numba evaluates it (0 when 0 > x, else x).
 * Otherwise as nx_erosion_run_tiled (rule 3 is exactly that call); results
 * equal the reference's FP64 step bit for bit. */
int nx_erosion_run_rule(nx_ctx *ctx, void *state, int64_t V,
                        const int32_t *col, const double *elen,
                        const nx_tile_plan *plan, int iters, double rain,
                        int rule, void *stream);

/* erosion.py:42-70 (erode_terrain1) over erosion.py:76-99
 * (erosion_iteration1): iters Jacobi steps in place on h (device f64 [V]),
 * each adding 0.0005 per higher neighbour and subtracting it per lower one
 * in slot order; tmp is a device scratch buffer of V doubles.  col as for
 * nx_erosion_run. */
int nx_erosion1_run(nx_ctx *ctx, double *h, double *tmp, int64_t V,
                    const int32_t *col, int iters, void *stream);

/* Store the current fields back as FP64 (any output may be NULL). */
int nx_erosion_store(nx_ctx *ctx, const void *state, int64_t V, double *h,
                     double *w, double *s, void *stream);

/* FP64 mass check: out (device f64[3]) = (sum(h + s), sum(|h| + |s|),
 * sum(w)).  In exact arithmetic every erosion step keeps h + s per vertex
 * (erosion.py:262-267), so out[0] is constant up to rounding. */
int nx_erosion_mass(nx_ctx *ctx, const void *state, int64_t V, double *out,
                    void *stream);

/* ---- batched planets (device) ------------------------------------------ */

/* One planet job of a batch: noise (terrain.py:32-59), optional rescale
 * (util.py:110-140) and optional erosion (erosion.py:171-192).  Host
 * struct; the pointers it holds are device pointers unless marked host. */
typedef struct nx_planet {
    const double *verts;     /* device f64 [V,3] */
    double *elev;            /* device f64 [V]: noise (then rescaled) heights;
                                eroded heights after nx_batch_store */
    int64_t V;
    int n_octaves;
    int accumulate;          /* add the octaves to elev's contents */
    double radius;           /* world_radius */
    const double *rough;     /* host f64 [n_octaves] */
    const double *strength;  /* host f64 [n_octaves] */
    const int32_t *perm;     /* host int32 [256] */
    const int32_t *pgi;      /* host int32 [256] */
    double lower, upper;     /* rescale bounds; lower >= upper: no rescale */
    void *erosion_state;     /* device, nx_erosion_state_bytes(V), or NULL */
    const int32_t *col;      /* device int32 [6,V] */
    const double *elen;      /* device f64 [6,V] */
    const nx_tile_plan *plan;/* host descriptor of a tile plan for col/elen
                                (its image on the device), or NULL */
    int64_t erosion_iters;   /* erosion steps of the whole job */
    double rain;             /* rain per step (reference: 0.3/320) */
    int64_t steps_done;      /* set to 0 by nx_batch_noise, advanced by
                                nx_batch_erode */
    int64_t step_cap;        /* nx_batch_erode: at most this many steps for
                                this planet in the call (0: max_steps) */
} nx_planet;

/* Workspace (device, 256-byte aligned) for a batch of n planets; one
 * workspace per stream (calls on that stream may reuse it). */
size_t nx_batch_workspace_bytes(int n);

/* Noise for every planet in one persistent launch (a device work queue of
 * `tile`-vertex items, `blocks_per_cu` blocks per CU), then per-planet
 * rescale and, for planets with an erosion state, the state load with the
 * first step's rain added.  Results equal nx_sample_octaves / nx_rescale /
 * nx_erosion_load on each planet alone. */
int nx_batch_noise(nx_ctx *ctx, nx_planet *planets, int n, void *workspace,
                   int tile, int blocks_per_cu, void *stream);

/* Up to `max_steps` erosion steps for every planet with steps left.
 * Planets with a tile plan take their steps as many per launch as the plan
 * serves (2 or 4; 2- and 4-step plans may be mixed): a 4-step planet first
 * takes a launch of 2 or 3 steps when its step count leaves one, and one
 * launch covers the tiles of all planets at the same point of the
 * schedule.  A last single step, and every step of planets without a plan,
 * runs one per launch (one launch covers all of them).  The sequence may be
 * replayed from a captured hipGraph.  Advances each planet's steps_done.
 * Results equal nx_erosion_run on each planet. */
int nx_batch_erode(nx_ctx *ctx, nx_planet *planets, int n, int max_steps,
                   void *workspace, int use_graph, void *stream);

/* Write the eroded heights of planets with an erosion state into elev. */
int nx_batch_store(nx_ctx *ctx, const nx_planet *planets, int n,
                   void *workspace, void *stream);

/* ---- mesh and neighbour tables (host) ---------------------------------- */

/* Counts of the icosphere with `d` divisions: V = 10 d^2 + 2, T = 20 d^2
 * (util.py:17-50 comment block). */
int nx_icosphere_counts(int64_t d, int64_t *V, int64_t *T);
/* util.py:17-50 (create_mesh, meshzoo.icosa_sphere): unit-sphere vertices
 * host f64 [V,3] (12 corner vertices first) and triangles host int32 [T,3]
 * (either may be NULL). */
int nx_icosphere(int64_t d, double *verts, int32_t *tris);
/* util.py:590-613 (build_adjacency): host int32 [V,6], -1 padded. */
int nx_build_adjacency(const int32_t *tris, int64_t T, int64_t V,
                       int32_t *adj);
/* util.py:638-662 (sort_adjacency): rows into cyclic order, in place. */
int nx_sort_adjacency(int32_t *adj, int64_t V);
/* ---- mesh and neighbour tables (device) -------------------------------- */

/* util.py:17-50: the nx_icosphere mesh generated on the device, equal to the
 * host one bit for bit.  verts device f64 [V,3], tris device int32 [T,3];
 * either may be NULL. */
int nx_icosphere_device(nx_ctx *ctx, int64_t d, double *verts, int32_t *tris,
                        void *stream);
/* util.py:590-613 on the device, for any triangle list: the same table as
 * nx_build_adjacency.  adj device int32 [V,6]; workspace device, 256-byte
 * aligned, of nx_adjacency_workspace_bytes(V) bytes.  A triangle index
 * outside 0..V-1 or a vertex in more than six triangles is reported by
 * nx_ctx_sync_check (NX_EINVAL). */
size_t nx_adjacency_workspace_bytes(int64_t V);
int nx_build_adjacency_device(nx_ctx *ctx, const int32_t *tris, int64_t T,
                              int64_t V, int32_t *adj, void *workspace,
                              void *stream);
/* util.py:638-662 on the device: adj_out = the rows of adj_in in cyclic
 * order (the same table as nx_sort_adjacency); the two must differ. */
int nx_sort_adjacency_device(nx_ctx *ctx, const int32_t *adj_in,
                             int32_t *adj_out, int64_t V, void *stream);
/* nx_sfc_order on the device (same order; verts device f64 [V,3], order
 * device int64 [V]), workspace of nx_sfc_workspace_bytes(V) bytes. */
size_t nx_sfc_workspace_bytes(int64_t V);
int nx_sfc_order_device(nx_ctx *ctx, const double *verts, int64_t V,
                        int64_t *order, void *workspace, void *stream);
/* nx_ell_from_adjacency on the device (same col and elen; all arrays on
 * the device).  With an order, workspace of nx_ell_workspace_bytes(V)
 * bytes; a bad neighbour index or order is reported by nx_ctx_sync_check. */
size_t nx_ell_workspace_bytes(int64_t V);
int nx_ell_from_adjacency_device(nx_ctx *ctx, const int32_t *adj, int64_t V,
                                 const double *verts, const int64_t *order,
                                 int32_t *col, double *elen, void *workspace,
                                 void *stream);

/* Space-filling-curve (3D Morton) vertex order: order[new] = old. */
int nx_sfc_order(const double *verts, int64_t V, int64_t *order);
/* Device neighbour layout from a [V,6] table (host arrays): col int32 [6,V]
 * and elen f64 [6,V]; order (nullable) renumbers vertices as in
 * nx_sfc_order.  Edge lengths follow erosion.py:38-40 in FP64. */
int nx_ell_from_adjacency(const int32_t *adj, int64_t V, const double *verts,
                          const int64_t *order, int32_t *col, double *elen);

#ifdef __cplusplus
}
#endif
#endif /* NIXIS_H */
