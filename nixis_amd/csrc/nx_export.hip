// nx_export.hip -- equirectangular export on the device (nixis.py:270-302,
// util.py:275-366): pixel coordinates (make_ll_arr), the k nearest mesh
// vertices of each pixel (the reference's scipy KD-tree query) and the
// inverse-distance sampling of make_gray_array.
//
// Nearest neighbours come from a uniform grid instead of a tree: points are
// sorted by the key of their cell, every occupied cell goes into an
// open-addressing hash table (key -> first point, count), and a query visits
// shells of cells around its own until the k-th best distance is no larger
// than the distance from the query to the faces of the block searched, so
// the result is the exact k nearest (ties ordered by point index).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "nx_ctx.h"
#include "nx_reduce_dev.h"

namespace nx {

namespace {

constexpr int kKnnMax = 8;                // largest k
constexpr unsigned long long kEmpty = ~0ull;
constexpr int kAxisBits = 21;             // cell coordinate bits per axis

struct GridHead {
    double origin[3];
    double cell;
    int32_t dims[3];
    int32_t pad;
    unsigned long long box[6];  // min/max keys per axis
};
static_assert(sizeof(GridHead) <= 256, "grid head must fit its 256 bytes");

struct CellEntry {
    unsigned long long key;
    int32_t start, count;
};

inline size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

uint64_t table_size(int64_t V)
{
    uint64_t t = 64;
    while (t < 2 * (uint64_t)V) t <<= 1;
    return t;
}

size_t sort_temp_bytes(int64_t V)
{
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint64_t *)nullptr,
                                             (uint64_t *)nullptr, (const int32_t *)nullptr,
                                             (int32_t *)nullptr, (int)V, 0, 3 * kAxisBits);
    return bytes;
}

struct GridLayout {
    size_t head, spts, sidx, kin, kout, iin, temp, table, total;
    uint64_t tsize;
};

GridLayout grid_layout(int64_t V)
{
    GridLayout L;
    size_t o = 0;
    L.head = o;
    o += 256;
    L.spts = o;
    o += a256(24 * (size_t)V);
    L.sidx = o;
    o += a256(4 * (size_t)V);
    L.kin = o;
    o += a256(8 * (size_t)V);
    L.kout = o;
    o += a256(8 * (size_t)V);
    L.iin = o;
    o += a256(4 * (size_t)V);
    L.temp = o;
    o += a256(sort_temp_bytes(V));
    L.tsize = table_size(V);
    L.table = o;
    o += a256(sizeof(CellEntry) * (size_t)L.tsize);
    L.total = o;
    return L;
}

// MurmurHash3's 64-bit finaliser.
__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

__device__ __forceinline__ unsigned long long cell_key(int x, int y, int z)
{
    return ((unsigned long long)x << (2 * kAxisBits)) | ((unsigned long long)y << kAxisBits) |
           (unsigned long long)z;
}

__device__ __forceinline__ int cell_of(double q, double o, double cell, int dim)
{
    const double t = floor((q - o) / cell);
    if (!(t >= 0.0)) return 0;  // below the grid, or NaN
    return t >= (double)(dim - 1) ? dim - 1 : (int)t;
}

}  // namespace

__global__ void k_grid_init(GridHead *g)
{
    if (threadIdx.x < 3) {
        g->box[2 * threadIdx.x] = ~0ull;
        g->box[2 * threadIdx.x + 1] = 0ull;
    }
}

__global__ __launch_bounds__(256) void k_grid_box(const double *__restrict__ pts, int64_t V,
                                                  GridHead *__restrict__ g)
{
    for (int k = 0; k < 3; ++k) {
        unsigned long long lo = ~0ull, hi = 0ull;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V;
             i += (int64_t)gridDim.x * blockDim.x) {
            const unsigned long long q = dkey(pts[3 * i + k]);
            lo = q < lo ? q : lo;
            hi = q > hi ? q : hi;
        }
        block_minmax(lo, hi);
        if (threadIdx.x == 0) {
            __hip_atomic_fetch_min(&g->box[2 * k], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_max(&g->box[2 * k + 1], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
}

// Grid geometry: a cell edge of twice the mean spacing of V points spread
// over the sphere that fits the box (about four to six points per occupied
// cell for a sphere mesh), at most 2^21 - 1 cells per axis.  The geometry
// only sets the speed; the search is exact for any cell size.
__global__ void k_grid_setup(GridHead *g, int64_t V)
{
    if (threadIdx.x != 0) return;
    double ext[3], lo[3];
    double R = 0.0;
    for (int k = 0; k < 3; ++k) {
        lo[k] = dunkey(g->box[2 * k]);
        ext[k] = dunkey(g->box[2 * k + 1]) - lo[k];
        R = fmax(R, 0.5 * ext[k]);
    }
    double cell = 2.0 * sqrt(4.0 * 3.141592653589793 * R * R / (double)V);
    for (int k = 0; k < 3; ++k) {
        const double need = ext[k] / 2097150.0;
        if (cell < need) cell = need;
    }
    if (!(cell > 0.0)) cell = 1.0;  // all points equal
    g->cell = cell;
    for (int k = 0; k < 3; ++k) {
        g->origin[k] = lo[k];
        const double n = floor(ext[k] / cell) + 1.0;
        g->dims[k] = (int32_t)fmin(n, 2097151.0);
    }
}

__global__ __launch_bounds__(256) void k_grid_keys(const double *__restrict__ pts, int64_t V,
                                                   const GridHead *__restrict__ g,
                                                   unsigned long long *__restrict__ key,
                                                   int32_t *__restrict__ idx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const double c = g->cell;
    const int x = cell_of(pts[3 * i], g->origin[0], c, g->dims[0]);
    const int y = cell_of(pts[3 * i + 1], g->origin[1], c, g->dims[1]);
    const int z = cell_of(pts[3 * i + 2], g->origin[2], c, g->dims[2]);
    key[i] = cell_key(x, y, z);
    idx[i] = (int32_t)i;
}

__global__ __launch_bounds__(256) void k_fill_cells(CellEntry *t, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        t[i].key = kEmpty;
        t[i].start = 0;
        t[i].count = 0;
    }
}

// Sorted coordinates, and one table entry per run of equal keys.  The table
// holds at least 2V slots, so every insertion finds a free one.
__global__ __launch_bounds__(256) void k_grid_fill(const double *__restrict__ pts, int64_t V,
                                                   const unsigned long long *__restrict__ skey,
                                                   const int32_t *__restrict__ sidx,
                                                   double *__restrict__ spts,
                                                   CellEntry *__restrict__ table, uint64_t mask)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= V) return;
    const int64_t src = sidx[j];
    spts[3 * j] = pts[3 * src];
    spts[3 * j + 1] = pts[3 * src + 1];
    spts[3 * j + 2] = pts[3 * src + 2];
    const unsigned long long k = skey[j];
    if (j > 0 && skey[j - 1] == k) return;
    int64_t e = j + 1;
    while (e < V && skey[e] == k) ++e;
    uint64_t slot = mix64(k) & mask;
    for (;;) {
        const unsigned long long prev = atomicCAS(&table[slot].key, kEmpty, k);
        if (prev == kEmpty) {
            table[slot].start = (int32_t)j;
            table[slot].count = (int32_t)(e - j);
            return;
        }
        slot = (slot + 1) & mask;
    }
}

namespace {

struct GridView {
    double ox, oy, oz, cell;
    int dx, dy, dz;
    const double *spts;
    const int32_t *sidx;
    const CellEntry *table;
    uint64_t mask;
};

__device__ __forceinline__ GridView grid_view(const char *ws, const double *spts,
                                              const int32_t *sidx, const CellEntry *table,
                                              uint64_t mask)
{
    const GridHead *g = (const GridHead *)ws;
    GridView v;
    v.ox = g->origin[0];
    v.oy = g->origin[1];
    v.oz = g->origin[2];
    v.cell = g->cell;
    v.dx = g->dims[0];
    v.dy = g->dims[1];
    v.dz = g->dims[2];
    v.spts = spts;
    v.sidx = sidx;
    v.table = table;
    v.mask = mask;
    return v;
}

__device__ __forceinline__ bool before(double s, int32_t i, double t, int32_t j)
{
    return s < t || (s == t && i < j);
}

template <int K>
__device__ __forceinline__ void visit_cell(const GridView &g, int x, int y, int z, double qx,
                                           double qy, double qz, double (&bd)[K],
                                           int32_t (&bi)[K])
{
    const unsigned long long key = cell_key(x, y, z);
    uint64_t slot = mix64(key) & g.mask;
    CellEntry e;
    for (;;) {
        e = g.table[slot];
        if (e.key == key) break;
        if (e.key == kEmpty) return;
        slot = (slot + 1) & g.mask;
    }
    for (int t = e.start; t < e.start + e.count; ++t) {
        double s = 0.0, d;
        d = qx - g.spts[3 * t];
        s += d * d;
        d = qy - g.spts[3 * t + 1];
        s += d * d;
        d = qz - g.spts[3 * t + 2];
        s += d * d;
        const int32_t id = g.sidx[t];
        if (!before(s, id, bd[K - 1], bi[K - 1])) continue;
        bd[K - 1] = s;
        bi[K - 1] = id;
#pragma unroll
        for (int j = K - 1; j > 0; --j)
            if (before(bd[j], bi[j], bd[j - 1], bi[j - 1])) {
                const double td = bd[j];
                bd[j] = bd[j - 1];
                bd[j - 1] = td;
                const int32_t ti = bi[j];
                bi[j] = bi[j - 1];
                bi[j - 1] = ti;
            }
    }
}

// The K nearest points of q: shells of cells around q's cell until the K-th
// squared distance is within the block's guaranteed radius (the distance
// from q to the nearest face of the block that does not lie on the grid's
// border, less a margin for rounding of the cell bounds).
template <int K>
__device__ void knn_point(const GridView &g, double qx, double qy, double qz, double (&bd)[K],
                          int32_t (&bi)[K])
{
#pragma unroll
    for (int j = 0; j < K; ++j) {
        bd[j] = INFINITY;
        bi[j] = INT32_MAX;
    }
    const int cx = cell_of(qx, g.ox, g.cell, g.dx);
    const int cy = cell_of(qy, g.oy, g.cell, g.dy);
    const int cz = cell_of(qz, g.oz, g.cell, g.dz);
    const int rmax = max(g.dx, max(g.dy, g.dz));
    const double margin = 1e-9 * g.cell;
    for (int r = 0; r <= rmax; ++r) {
        for (int z = cz - r; z <= cz + r; ++z) {
            if (z < 0 || z >= g.dz) continue;
            for (int y = cy - r; y <= cy + r; ++y) {
                if (y < 0 || y >= g.dy) continue;
                const bool face = (z == cz - r || z == cz + r || y == cy - r || y == cy + r);
                const int step = (face || r == 0) ? 1 : 2 * r;
                for (int x = cx - r; x <= cx + r; x += step) {
                    if (x < 0 || x >= g.dx) continue;
                    visit_cell<K>(g, x, y, z, qx, qy, qz, bd, bi);
                }
            }
        }
        double guar = INFINITY;
        if (cx - r > 0) guar = fmin(guar, qx - (g.ox + (cx - r) * g.cell));
        if (cx + r + 1 < g.dx) guar = fmin(guar, (g.ox + (cx + r + 1) * g.cell) - qx);
        if (cy - r > 0) guar = fmin(guar, qy - (g.oy + (cy - r) * g.cell));
        if (cy + r + 1 < g.dy) guar = fmin(guar, (g.oy + (cy + r + 1) * g.cell) - qy);
        if (cz - r > 0) guar = fmin(guar, qz - (g.oz + (cz - r) * g.cell));
        if (cz + r + 1 < g.dz) guar = fmin(guar, (g.oz + (cz + r + 1) * g.cell) - qz);
        if (guar == INFINITY) break;  // the block covers the whole grid
        guar -= margin;
        if (guar > 0.0 && bd[K - 1] <= guar * guar) break;
    }
}

template <int K>
__device__ __forceinline__ void knn_store(int64_t p, const double (&bd)[K], const int32_t (&bi)[K],
                                          int64_t V, double *dists, int64_t *nbrs)
{
#pragma unroll
    for (int j = 0; j < K; ++j) {
        dists[p * K + j] = sqrt(bd[j]);  // +inf where fewer than K points exist
        nbrs[p * K + j] = bi[j] == INT32_MAX ? V : bi[j];
    }
}

}  // namespace

template <int K>
__global__ __launch_bounds__(256) void k_knn_points(const char *__restrict__ ws,
                                                    const double *__restrict__ spts,
                                                    const int32_t *__restrict__ sidx,
                                                    const CellEntry *__restrict__ table,
                                                    uint64_t mask, int64_t V,
                                                    const double *__restrict__ q, int64_t nq,
                                                    double *__restrict__ dists,
                                                    int64_t *__restrict__ nbrs)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nq) return;
    const GridView g = grid_view(ws, spts, sidx, table, mask);
    double bd[K];
    int32_t bi[K];
    knn_point<K>(g, q[3 * p], q[3 * p + 1], q[3 * p + 2], bd, bi);
    knn_store<K>(p, bd, bi, V, dists, nbrs);
}

// trig: cos(lat_h), sin(lat_h) for each row, then cos(lon_w), sin(lon_w) for
// each column, computed on the host with the C library the reference's
// numba code calls; x = (r * cos lat) * cos lon, y = (r * cos lat) * sin lon,
// z = r * sin lat as util.py:86-88.
template <int K>
__global__ __launch_bounds__(256) void k_knn_pixels(const char *__restrict__ ws,
                                                    const double *__restrict__ spts,
                                                    const int32_t *__restrict__ sidx,
                                                    const CellEntry *__restrict__ table,
                                                    uint64_t mask, int64_t V, int64_t width,
                                                    int64_t height, double radius,
                                                    const double *__restrict__ trig,
                                                    double *__restrict__ dists,
                                                    int64_t *__restrict__ nbrs)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= width * height) return;
    const int64_t h = p / width, w = p - h * width;
    const double clat = trig[2 * h], slat = trig[2 * h + 1];
    const double clon = trig[2 * height + 2 * w], slon = trig[2 * height + 2 * w + 1];
    const double qx = radius * clat * clon, qy = radius * clat * slon, qz = radius * slat;
    const GridView g = grid_view(ws, spts, sidx, table, mask);
    double bd[K];
    int32_t bi[K];
    knn_point<K>(g, qx, qy, qz, bd, bi);
    knn_store<K>(p, bd, bi, V, dists, nbrs);
}

__global__ __launch_bounds__(256) void k_ll_arr(int64_t width, int64_t height, double radius,
                                                const double *__restrict__ trig,
                                                double *__restrict__ out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= width * height) return;
    const int64_t h = p / width, w = p - h * width;
    const double clat = trig[2 * h], slat = trig[2 * h + 1];
    const double clon = trig[2 * height + 2 * w], slon = trig[2 * height + 2 * w + 1];
    out[3 * p] = radius * clat * clon;
    out[3 * p + 1] = radius * clat * slon;
    out[3 * p + 2] = radius * slat;
}

// util.py:356-364 over k neighbours: sd = sum(d); w_i = 1 / ((d_i + 0.00001)
// / sd); t = sum(w); value = int(sum(colors[n_i] * (w_i / t))), each sum
// accumulated from 0 in neighbour order; truncation toward zero.
__global__ __launch_bounds__(256) void k_gray(int64_t npix, int k, const double *__restrict__ dists,
                                              const int64_t *__restrict__ nbrs,
                                              const double *__restrict__ colors, int64_t ncolors,
                                              int32_t *__restrict__ out, int *__restrict__ err)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    double d[kKnnMax], ws[kKnnMax];
    double sd = 0.0;
    for (int i = 0; i < k; ++i) {
        d[i] = dists[p * k + i];
        sd += d[i];
    }
    double t = 0.0;
    for (int i = 0; i < k; ++i) {
        ws[i] = 1.0 / ((d[i] + 0.00001) / sd);
        t += ws[i];
    }
    double v = 0.0;
    bool bad = false;
    for (int i = 0; i < k; ++i) {
        const int64_t n = nbrs[p * k + i];
        if (n < 0 || n >= ncolors) {
            bad = true;
            continue;
        }
        v += colors[n] * (ws[i] / t);
    }
    out[p] = (int32_t)(int64_t)v;
    if (bad) __hip_atomic_fetch_or(err, NX_ERR_INDEX_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

size_t knn_grid_bytes(int64_t V)
{
    if (V <= 0 || V >= INT32_MAX) return 256;
    return grid_layout(V).total;
}

int knn_grid_build(nx_ctx *ctx, const double *pts, int64_t V, void *grid, hipStream_t stream)
{
    if (!ctx || !pts || !grid || V <= 0 || V >= INT32_MAX)
        return set_error(NX_EINVAL, "grid build needs points, a workspace and 0 < V < 2^31 - 1");
    if (((uintptr_t)grid & 255) != 0) return set_error(NX_EINVAL, "grid workspace must be 256-byte aligned");
    NX_HIP(hipSetDevice(ctx->device));
    const GridLayout L = grid_layout(V);
    char *w = (char *)grid;
    auto *g = (GridHead *)(w + L.head);
    auto *kin = (unsigned long long *)(w + L.kin);
    auto *kout = (unsigned long long *)(w + L.kout);
    auto *iin = (int32_t *)(w + L.iin);
    auto *sidx = (int32_t *)(w + L.sidx);
    auto *spts = (double *)(w + L.spts);
    auto *table = (CellEntry *)(w + L.table);
    const unsigned blocks = (unsigned)((V + 255) / 256);
    hipLaunchKernelGGL(k_grid_init, dim3(1), dim3(64), 0, stream, g);
    hipLaunchKernelGGL(k_grid_box, dim3(std::min<unsigned>(blocks, 2048u)), dim3(256), 0, stream,
                       pts, V, g);
    hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(64), 0, stream, g, V);
    hipLaunchKernelGGL(k_grid_keys, dim3(blocks), dim3(256), 0, stream, pts, V, g, kin, iin);
    size_t temp = sort_temp_bytes(V);
    NX_HIP(hipcub::DeviceRadixSort::SortPairs(w + L.temp, temp, (uint64_t *)kin, (uint64_t *)kout,
                                              iin, sidx, (int)V, 0, 3 * kAxisBits, stream));
    hipLaunchKernelGGL(k_fill_cells, dim3((unsigned)((L.tsize + 255) / 256)), dim3(256), 0, stream,
                       table, L.tsize);
    hipLaunchKernelGGL(k_grid_fill, dim3(blocks), dim3(256), 0, stream, pts, V, kout, sidx, spts,
                       table, L.tsize - 1);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

template <int K>
static void launch_points(const char *w, const GridLayout &L, int64_t V, const double *q,
                          int64_t nq, double *dists, int64_t *nbrs, hipStream_t stream)
{
    hipLaunchKernelGGL(k_knn_points<K>, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, stream,
                       w, (const double *)(w + L.spts), (const int32_t *)(w + L.sidx),
                       (const CellEntry *)(w + L.table), L.tsize - 1, V, q, nq, dists, nbrs);
}

template <int K>
static void launch_pixels(const char *w, const GridLayout &L, int64_t V, int64_t width,
                          int64_t height, double radius, const double *trig, double *dists,
                          int64_t *nbrs, hipStream_t stream)
{
    const int64_t n = width * height;
    hipLaunchKernelGGL(k_knn_pixels<K>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       w, (const double *)(w + L.spts), (const int32_t *)(w + L.sidx),
                       (const CellEntry *)(w + L.table), L.tsize - 1, V, width, height, radius,
                       trig, dists, nbrs);
}

#define NX_K_DISPATCH(FN, ...)                  \
    switch (k) {                                \
    case 1: FN<1>(__VA_ARGS__); break;          \
    case 2: FN<2>(__VA_ARGS__); break;          \
    case 3: FN<3>(__VA_ARGS__); break;          \
    case 4: FN<4>(__VA_ARGS__); break;          \
    case 5: FN<5>(__VA_ARGS__); break;          \
    case 6: FN<6>(__VA_ARGS__); break;          \
    case 7: FN<7>(__VA_ARGS__); break;          \
    default: FN<8>(__VA_ARGS__); break;         \
    }

int knn_query(nx_ctx *ctx, const void *grid, int64_t V, const double *q, int64_t nq, int k,
              double *dists, int64_t *nbrs, hipStream_t stream)
{
    if (!ctx || !grid || V <= 0 || V >= INT32_MAX || nq < 0 || (nq > 0 && (!q || !dists || !nbrs)))
        return set_error(NX_EINVAL, "query needs a built grid, query points and outputs");
    if (k < 1 || k > kKnnMax) return set_error(NX_EINVAL, "k=%d outside 1..%d", k, kKnnMax);
    if (nq == 0) return NX_OK;
    if (nq > (int64_t)UINT32_MAX * 256) return set_error(NX_EINVAL, "too many query points");
    NX_HIP(hipSetDevice(ctx->device));
    const GridLayout L = grid_layout(V);
    NX_K_DISPATCH(launch_points, (const char *)grid, L, V, q, nq, dists, nbrs, stream);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

size_t pixel_workspace_bytes(int64_t width, int64_t height)
{
    return width > 0 && height > 0 ? a256(16 * (size_t)(width + height)) : 256;
}

// cos/sin of each row's latitude and each column's longitude in radians, as
// util.py:300-309 computes lat, lon and latlon2xyz its arguments.  One
// sincos per angle: compilers pair a sin and a cos of the same operand into
// the C library's sincos (LLVM for numba's intrinsics, GCC for C), and
// glibc's sincos differs from separate sin and cos in the last bit for a
// few arguments.
static int stage_trig(nx_ctx *ctx, int64_t width, int64_t height, void *ws, hipStream_t stream)
{
    const size_t bytes = 16 * (size_t)(width + height);
    int slot = 0;
    double *t = (double *)stage_begin(ctx, bytes, &slot);
    if (!t) return set_error(NX_ENOMEM, "pinned staging buffer allocation failed");
    const double latpercent = 180.0 / (double)(height - 1);
    const double lonpercent = 360.0 / (double)width;
    const double d2r = 3.141592653589793 / 180.0;  // np.pi / 180
    for (int64_t h = 0; h < height; ++h) {
        double lat = 90.0 - ((double)h * latpercent);
        if (lat < -90.0) lat = -90.0;
        sincos(lat * d2r, &t[2 * h + 1], &t[2 * h]);
    }
    for (int64_t w = 0; w < width; ++w) {
        double lon = -180.0 + ((double)w * lonpercent);
        if (lon < -180.0) lon = -180.0;
        sincos(lon * d2r, &t[2 * height + 2 * w + 1], &t[2 * height + 2 * w]);
    }
    return stage_commit(ctx, slot, ws, bytes, stream);
}

static int pixel_check(nx_ctx *ctx, int64_t width, int64_t height, const void *ws)
{
    if (!ctx || !ws) return set_error(NX_EINVAL, "pixel grid needs a context and a workspace");
    if (width < 1 || height < 2 || width > (1 << 24) || height > (1 << 24) ||
        width * height > (int64_t)UINT32_MAX * 256)
        return set_error(NX_EINVAL, "image %lld x %lld outside 1..2^24 x 2..2^24",
                         (long long)width, (long long)height);
    return NX_OK;
}

int knn_query_pixels(nx_ctx *ctx, const void *grid, int64_t V, int64_t width, int64_t height,
                     double radius, int k, double *dists, int64_t *nbrs, void *ws,
                     hipStream_t stream)
{
    int rc = pixel_check(ctx, width, height, ws);
    if (rc) return rc;
    if (!grid || V <= 0 || V >= INT32_MAX || !dists || !nbrs)
        return set_error(NX_EINVAL, "query needs a built grid and outputs");
    if (k < 1 || k > kKnnMax) return set_error(NX_EINVAL, "k=%d outside 1..%d", k, kKnnMax);
    NX_HIP(hipSetDevice(ctx->device));
    if ((rc = stage_trig(ctx, width, height, ws, stream))) return rc;
    const GridLayout L = grid_layout(V);
    NX_K_DISPATCH(launch_pixels, (const char *)grid, L, V, width, height, radius,
                  (const double *)ws, dists, nbrs, stream);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

int make_ll_arr(nx_ctx *ctx, int64_t width, int64_t height, double radius, double *out, void *ws,
                hipStream_t stream)
{
    int rc = pixel_check(ctx, width, height, ws);
    if (rc) return rc;
    if (!out) return set_error(NX_EINVAL, "null output");
    NX_HIP(hipSetDevice(ctx->device));
    if ((rc = stage_trig(ctx, width, height, ws, stream))) return rc;
    const int64_t n = width * height;
    hipLaunchKernelGGL(k_ll_arr, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, width,
                       height, radius, (const double *)ws, out);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

int make_gray_array(nx_ctx *ctx, int64_t npix, int k, const double *dists, const int64_t *nbrs,
                    const double *colors, int64_t ncolors, int32_t *out, hipStream_t stream)
{
    if (!ctx || npix < 0 || (npix > 0 && (!dists || !nbrs || !colors || !out)) || ncolors <= 0)
        return set_error(NX_EINVAL, "gray array needs query results, colors and an output");
    if (k < 1 || k > kKnnMax) return set_error(NX_EINVAL, "k=%d outside 1..%d", k, kKnnMax);
    if (npix == 0) return NX_OK;
    if (npix > (int64_t)UINT32_MAX * 256) return set_error(NX_EINVAL, "too many pixels");
    NX_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_gray, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, stream, npix, k,
                       dists, nbrs, colors, ncolors, out, ctx->d_err);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

}  // namespace nx
