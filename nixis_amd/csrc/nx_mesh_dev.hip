// nx_mesh_dev.hip -- icosphere and neighbour tables generated on the device.
//
// The same mesh and tables as the host code in nx_mesh.cpp, bit for bit
// (same FP64 operations, built without contraction), for the reference's
// mesh step (util.py:17-50, meshzoo) and its table builders (util.py:590-662):
//   * k_ico_points: one thread per vertex; corners, then edge points, then
//     face interiors row by row, each the projection onto the unit sphere of
//     the same weighted sum of corner vectors as on the host;
//   * k_ico_tris: one thread per triangle, rows of "up" then "down"
//     triangles per face, counter-clockwise;
//   * build_adjacency (util.py:590-613) appends, for every triangle in
//     order, its next vertex to each corner's row.  Here every (vertex,
//     triangle) pair takes a slot with an atomic counter, then each row is
//     put back into triangle order (at most six entries), which is the
//     reference's append order;
//   * sort_adjacency (util.py:638-662) walks each vertex's ring from its
//     first entry, reading the unsorted table (what a race-free run of the
//     reference computes) and writing a second table.
#include <hip/hip_runtime.h>

#include <cmath>

#include "nx_ctx.h"

namespace nx {

namespace {

// Icosahedron (edge-up) corners and counter-clockwise faces, as nx_mesh.cpp.
constexpr double kPhi = 1.6180339887498948482;
constexpr double kIcoV[12][3] = {
    {-1, kPhi, 0}, {1, kPhi, 0},  {-1, -kPhi, 0}, {1, -kPhi, 0},
    {0, -1, kPhi}, {0, 1, kPhi},  {0, -1, -kPhi}, {0, 1, -kPhi},
    {kPhi, 0, -1}, {kPhi, 0, 1},  {-kPhi, 0, -1}, {-kPhi, 0, 1},
};
constexpr int kIcoF[20][3] = {
    {0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
    {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
    {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
    {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1},
};

// Everything the kernels need about the base icosahedron.
struct IcoArgs {
    double c[12][3];     // unit corner vectors
    int8_t face[20][3];
    int8_t eid[12][12];  // edge id of a corner pair, -1 if none
    int8_t ea[30], eb[30];
    int64_t d;
};

IcoArgs make_args(int64_t d)
{
    IcoArgs a;
    a.d = d;
    for (int i = 0; i < 12; ++i) {
        const double x = kIcoV[i][0], y = kIcoV[i][1], z = kIcoV[i][2];
        const double n = std::sqrt(x * x + y * y + z * z);
        a.c[i][0] = x / n;
        a.c[i][1] = y / n;
        a.c[i][2] = z / n;
    }
    for (auto &r : a.eid)
        for (auto &x : r) x = -1;
    int n = 0;
    for (int f = 0; f < 20; ++f)
        for (int k = 0; k < 3; ++k) {
            a.face[f][k] = (int8_t)kIcoF[f][k];
            const int u = kIcoF[f][k], v = kIcoF[f][(k + 1) % 3];
            const int lo = u < v ? u : v, hi = u < v ? v : u;
            if (a.eid[lo][hi] < 0) {
                a.eid[lo][hi] = a.eid[hi][lo] = (int8_t)n;
                a.ea[n] = (int8_t)lo;
                a.eb[n] = (int8_t)hi;
                ++n;
            }
        }
    return a;
}

__device__ __forceinline__ void project(double x, double y, double z, double *out)
{
    const double n = sqrt(x * x + y * y + z * z);
    out[0] = x / n;
    out[1] = y / n;
    out[2] = z / n;
}

__device__ __forceinline__ int64_t edge_point(const IcoArgs &g, int u, int v, int64_t k)
{
    const int id = g.eid[u][v];
    const int64_t kk = (u < v) ? k : g.d - k;
    return 12 + (int64_t)id * (g.d - 1) + (kk - 1);
}

// Global index of face point (r, c), 0 <= c <= r <= d (nx_mesh.cpp Grid::index).
__device__ __forceinline__ int64_t grid_index(const IcoArgs &g, int f, int64_t r, int64_t c)
{
    const int A = g.face[f][0], B = g.face[f][1], C = g.face[f][2];
    const int64_t d = g.d;
    if (r == 0) return A;
    if (r == d && c == 0) return B;
    if (r == d && c == d) return C;
    if (c == 0) return edge_point(g, A, B, r);
    if (c == r) return edge_point(g, A, C, r);
    if (r == d) return edge_point(g, B, C, c);
    const int64_t base = 12 + 30 * (d - 1) + (int64_t)f * ((d - 1) * (d - 2) / 2);
    return base + (r - 1) * (r - 2) / 2 + (c - 1);
}

// Largest r >= 0 with r * r <= x.
__device__ __forceinline__ int64_t isqrt64(int64_t x)
{
    int64_t r = (int64_t)sqrt((double)x);
    while (r * r > x) --r;
    while ((r + 1) * (r + 1) <= x) ++r;
    return r;
}

}  // namespace

__global__ __launch_bounds__(256) void k_ico_points(IcoArgs g, int64_t V, double *__restrict__ verts)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const int64_t d = g.d;
    double *o = verts + 3 * i;
    if (i < 12) {
        o[0] = g.c[i][0];
        o[1] = g.c[i][1];
        o[2] = g.c[i][2];
        return;
    }
    const int64_t ne = 30 * (d - 1);
    if (i < 12 + ne) {
        // edge point: ((d-k) U + k W) projected, U the lower-numbered corner
        const int64_t j = i - 12;
        const int id = (int)(j / (d - 1));
        const int64_t k = j % (d - 1) + 1;
        const double *U = g.c[g.ea[id]], *W = g.c[g.eb[id]];
        const double p = (double)(d - k), q = (double)k;
        project(p * U[0] + q * W[0], p * U[1] + q * W[1], p * U[2] + q * W[2], o);
        return;
    }
    // interior point (r, c), 2 <= r <= d-1, 1 <= c <= r-1: local index
    // l = (r-1)(r-2)/2 + (c-1) within its face
    const int64_t m = (d - 1) * (d - 2) / 2;
    const int64_t j = i - 12 - ne;
    const int f = (int)(j / m);
    const int64_t l = j % m;
    // s = r - 1 is the largest s with s (s - 1) / 2 <= l
    int64_t s = (int64_t)((1.0 + sqrt(1.0 + 8.0 * (double)l)) * 0.5);
    while (s * (s - 1) / 2 > l) --s;
    while ((s + 1) * s / 2 <= l) ++s;
    const int64_t r = s + 1, c = l - (r - 1) * (r - 2) / 2 + 1;
    const double *A = g.c[g.face[f][0]], *B = g.c[g.face[f][1]], *C = g.c[g.face[f][2]];
    const double a = (double)(d - r), bb = (double)(r - c), w = (double)c;
    project(a * A[0] + bb * B[0] + w * C[0], a * A[1] + bb * B[1] + w * C[1],
            a * A[2] + bb * B[2] + w * C[2], o);
}

__global__ __launch_bounds__(256) void k_ico_tris(IcoArgs g, int64_t T, int32_t *__restrict__ tris)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const int64_t d2 = g.d * g.d;
    const int f = (int)(t / d2);
    const int64_t tl = t % d2;
    const int64_t r = isqrt64(tl);  // rows before r hold r^2 triangles
    const int64_t off = tl - r * r;
    int32_t *o = tris + 3 * t;
    if (off <= r) {  // "up" triangle c = off
        o[0] = (int32_t)grid_index(g, f, r, off);
        o[1] = (int32_t)grid_index(g, f, r + 1, off);
        o[2] = (int32_t)grid_index(g, f, r + 1, off + 1);
    } else {  // "down" triangle c = off - r - 1
        const int64_t c = off - r - 1;
        o[0] = (int32_t)grid_index(g, f, r, c);
        o[1] = (int32_t)grid_index(g, f, r + 1, c + 1);
        o[2] = (int32_t)grid_index(g, f, r, c + 1);
    }
}

// build_adjacency, pass 1: every triangle corner appends (triangle, next
// vertex) to its vertex's slots; a seventh entry marks the vertex bad.
__global__ __launch_bounds__(256) void k_adj_append(const int32_t *__restrict__ tris, int64_t T,
                                                    int64_t V, int32_t *__restrict__ cnt,
                                                    int32_t *__restrict__ tri_of,
                                                    int32_t *__restrict__ adj, int *__restrict__ err)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const int32_t v[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
    for (int k = 0; k < 3; ++k) {
        const int32_t a = v[k], b = v[(k + 1) % 3];
        if (a < 0 || a >= V || b < 0 || b >= V) {
            __hip_atomic_fetch_or(err, NX_ERR_MESH_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        const int slot = atomicAdd(&cnt[a], 1);
        if (slot >= 6) {
            __hip_atomic_fetch_or(err, NX_ERR_MESH_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        tri_of[(int64_t)a * 6 + slot] = (int32_t)t;
        adj[(int64_t)a * 6 + slot] = b;
    }
}

// build_adjacency, pass 2: each row back into triangle order, -1 padded.
__global__ __launch_bounds__(256) void k_adj_order(int64_t V, const int32_t *__restrict__ cnt,
                                                   const int32_t *__restrict__ tri_of,
                                                   int32_t *__restrict__ adj)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int n = cnt[v] < 6 ? cnt[v] : 6;
    int32_t t[6], b[6];
    for (int k = 0; k < 6; ++k) {
        t[k] = k < n ? tri_of[v * 6 + k] : 0x7fffffff;
        b[k] = k < n ? adj[v * 6 + k] : -1;
    }
    for (int i = 1; i < 6; ++i)  // insertion sort by triangle index
        for (int k = i; k > 0 && t[k - 1] > t[k]; --k) {
            const int32_t tt = t[k]; t[k] = t[k - 1]; t[k - 1] = tt;
            const int32_t bb = b[k]; b[k] = b[k - 1]; b[k - 1] = bb;
        }
    for (int k = 0; k < 6; ++k) adj[v * 6 + k] = b[k];
}

// sort_adjacency (util.py:638-662, next_vert util.py:623-629).
__global__ __launch_bounds__(256) void k_adj_sort(int64_t V, const int32_t *__restrict__ in,
                                                  int32_t *__restrict__ out, int *__restrict__ err)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= V) return;
    const int deg = idx < 12 ? 5 : 6;
    int32_t row[6];
    for (int k = 0; k < 6; ++k) row[k] = in[idx * 6 + k];
    int32_t visited[6] = {-1, -1, -1, -1, -1, -1};
    int32_t pv = (int32_t)idx, nv = row[0];
    bool bad = false;
    for (int i = 0; i < deg - 1; ++i) {
        visited[i] = nv;
        int32_t next = -1;
        if (nv >= 0 && nv < V) {
            int32_t other[6];
            for (int k = 0; k < 6; ++k) other[k] = in[(int64_t)nv * 6 + k];
            for (int a = 0; a < 6 && next < 0; ++a) {
                const int32_t cand = row[a];
                for (int b = 0; b < 6; ++b)
                    if (other[b] == cand) {
                        if (cand != pv) next = cand;
                        break;
                    }
            }
        }
        nv = next;
        pv = visited[i];
    }
    visited[deg - 1] = nv;
    for (int k = 0; k < 6; ++k) {
        out[idx * 6 + k] = visited[k];
        bad = bad | (k < deg && visited[k] < 0);
    }
    if (bad) __hip_atomic_fetch_or(err, NX_ERR_MESH_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_zero_i32(int32_t *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

static unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

int icosphere_dev(nx_ctx *ctx, int64_t d, double *verts, int32_t *tris, hipStream_t stream)
{
    if (!ctx) return set_error(NX_EINVAL, "null context");
    int64_t V = 0, T = 0;
    if (icosphere_counts(d, &V, &T) != NX_OK || V > INT32_MAX)
        return set_error(NX_EINVAL, "divisions %lld: V must fit int32", (long long)d);
    NX_HIP(hipSetDevice(ctx->device));
    const IcoArgs g = make_args(d);
    if (verts) hipLaunchKernelGGL(k_ico_points, dim3(blocks_for(V)), dim3(256), 0, stream, g, V, verts);
    if (tris) hipLaunchKernelGGL(k_ico_tris, dim3(blocks_for(T)), dim3(256), 0, stream, g, T, tris);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

size_t adjacency_workspace_bytes(int64_t V) { return V > 0 ? (size_t)V * 7 * sizeof(int32_t) + 256 : 256; }

int build_adjacency_dev(nx_ctx *ctx, const int32_t *tris, int64_t T, int64_t V, int32_t *adj,
                        void *workspace, hipStream_t stream)
{
    if (!ctx || !tris || !adj || !workspace || T < 0 || V <= 0 || V > INT32_MAX || T > INT32_MAX)
        return set_error(NX_EINVAL, "build_adjacency needs triangles, V in 1..2^31-1, a table and a workspace");
    if (((uintptr_t)workspace & 255) != 0) return set_error(NX_EINVAL, "workspace must be 256-byte aligned");
    NX_HIP(hipSetDevice(ctx->device));
    int32_t *cnt = (int32_t *)workspace;
    int32_t *tri_of = cnt + ((V + 63) / 64) * 64;
    hipLaunchKernelGGL(k_zero_i32, dim3(blocks_for(V)), dim3(256), 0, stream, cnt, V);
    if (T > 0)
        hipLaunchKernelGGL(k_adj_append, dim3(blocks_for(T)), dim3(256), 0, stream, tris, T, V, cnt,
                           tri_of, adj, ctx->d_err);
    hipLaunchKernelGGL(k_adj_order, dim3(blocks_for(V)), dim3(256), 0, stream, V, cnt, tri_of, adj);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

int sort_adjacency_dev(nx_ctx *ctx, const int32_t *adj_in, int32_t *adj_out, int64_t V,
                       hipStream_t stream)
{
    if (!ctx || !adj_in || !adj_out || V < 12 || V > INT32_MAX || adj_in == adj_out)
        return set_error(NX_EINVAL, "sort_adjacency needs two distinct tables and V >= 12");
    NX_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_adj_sort, dim3(blocks_for(V)), dim3(256), 0, stream, V, adj_in, adj_out,
                       ctx->d_err);
    NX_HIP(hipGetLastError());
    return NX_OK;
}

}  // namespace nx
