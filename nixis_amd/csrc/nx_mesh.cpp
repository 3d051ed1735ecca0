// nx_mesh.cpp -- host-side icosphere and neighbour-table construction.
//
// The reference takes its mesh from meshzoo.icosa_sphere (util.py:17-50),
// then builds a [V,6] neighbour table with build_adjacency (util.py:590-613)
// and puts each row into cyclic order with sort_adjacency (util.py:638-662).
// meshzoo is an external package, so this file defines its own icosphere with
// the same counts (V = 10d^2+2, T = 20d^2), the same corner-first vertex
// numbering (the 12 degree-5 vertices are 0..11, which sort_adjacency relies
// on) and flat subdivision projected onto the unit sphere.
//
// The neighbour table is then converted into the device layout used by the
// erosion kernel: slot-major ("ELL") arrays col[6][V] and elen[6][V], with the
// 12 pentagons' sixth slot pointing at the vertex itself (a self-neighbour
// compares equal in height, so it contributes nothing, exactly like the
// reference skipping -1).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "nx_internal.h"

namespace {

// Icosahedron with an edge centred on each pole (edge-up).
const double PHI = 1.6180339887498948482;
const double ICO_V[12][3] = {
    {-1, PHI, 0}, {1, PHI, 0},  {-1, -PHI, 0}, {1, -PHI, 0},
    {0, -1, PHI}, {0, 1, PHI},  {0, -1, -PHI}, {0, 1, -PHI},
    {PHI, 0, -1}, {PHI, 0, 1},  {-PHI, 0, -1}, {-PHI, 0, 1},
};
// Counter-clockwise seen from outside.
const int ICO_F[20][3] = {
    {0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
    {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
    {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
    {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1},
};

struct Edges {
    int id[12][12];  // edge id for an unordered corner pair, -1 if none
    int a[30], b[30];
};

Edges make_edges()
{
    Edges e;
    for (auto &r : e.id)
        for (int &x : r) x = -1;
    int n = 0;
    for (int f = 0; f < 20; ++f)
        for (int k = 0; k < 3; ++k) {
            int u = ICO_F[f][k], v = ICO_F[f][(k + 1) % 3];
            int lo = std::min(u, v), hi = std::max(u, v);
            if (e.id[lo][hi] < 0) {
                e.id[lo][hi] = e.id[hi][lo] = n;
                e.a[n] = lo;
                e.b[n] = hi;
                ++n;
            }
        }
    return e;
}

// Unit corner vectors.
void corners(double c[12][3])
{
    for (int i = 0; i < 12; ++i) {
        double x = ICO_V[i][0], y = ICO_V[i][1], z = ICO_V[i][2];
        double n = std::sqrt(x * x + y * y + z * z);
        c[i][0] = x / n;
        c[i][1] = y / n;
        c[i][2] = z / n;
    }
}

inline void project(double x, double y, double z, double *out)
{
    double n = std::sqrt(x * x + y * y + z * z);
    out[0] = x / n;
    out[1] = y / n;
    out[2] = z / n;
}

struct Grid {
    int64_t d;
    Edges e;
    // Global index of face point (r, c), 0 <= c <= r <= d, barycentric
    // (d-r, r-c, c) on the face's corners (A, B, C).
    int64_t index(int f, int64_t r, int64_t c) const
    {
        const int A = ICO_F[f][0], B = ICO_F[f][1], C = ICO_F[f][2];
        if (r == 0) return A;
        if (r == d && c == 0) return B;
        if (r == d && c == d) return C;
        if (c == 0) return edge_point(A, B, r);          // on AB, r steps from A
        if (c == r) return edge_point(A, C, r);          // on AC, r steps from A
        if (r == d) return edge_point(B, C, c);          // on BC, c steps from B
        const int64_t base = 12 + 30 * (d - 1) + (int64_t)f * ((d - 1) * (d - 2) / 2);
        return base + (r - 1) * (r - 2) / 2 + (c - 1);
    }
    // Point k steps (1..d-1) from corner u toward corner v.
    int64_t edge_point(int u, int v, int64_t k) const
    {
        const int id = e.id[u][v];
        const int64_t kk = (u < v) ? k : d - k;  // canonical: from lower corner
        return 12 + (int64_t)id * (d - 1) + (kk - 1);
    }
};

}  // namespace

namespace nx {

int icosphere_counts(int64_t d, int64_t *V, int64_t *T)
{
    if (d < 1 || d > 100000) return NX_EINVAL;
    if (V) *V = 10 * d * d + 2;
    if (T) *T = 20 * d * d;
    return NX_OK;
}

int icosphere(int64_t d, double *verts, int32_t *tris)
{
    if (d < 1 || d > 100000) return NX_EINVAL;
    const int64_t V = 10 * d * d + 2;
    if (V > INT32_MAX) return NX_EINVAL;
    Grid g{d, make_edges()};
    double c[12][3];
    corners(c);
    if (verts) {
        for (int i = 0; i < 12; ++i)
            for (int k = 0; k < 3; ++k) verts[3 * i + k] = c[i][k];
        // Edge points: ((d-k) U + k V) projected, U the lower-numbered corner.
        parallel_for(30 * (d - 1), [&](int64_t b, int64_t e) {
            for (int64_t j = b; j < e; ++j) {
                const int id = (int)(j / (d - 1));
                const int64_t k = j % (d - 1) + 1;
                const double *U = c[g.e.a[id]], *W = c[g.e.b[id]];
                const double p = (double)(d - k), q = (double)k;
                project(p * U[0] + q * W[0], p * U[1] + q * W[1], p * U[2] + q * W[2],
                        verts + 3 * (12 + j));
            }
        }, 1024);
        // Interior points: barycentric (d-r, r-c, c) projected.
        parallel_for(20 * d, [&](int64_t b, int64_t e) {
            for (int64_t j = b; j < e; ++j) {
                const int f = (int)(j / d);
                const int64_t r = j % d;
                const double *A = c[ICO_F[f][0]], *B = c[ICO_F[f][1]], *C = c[ICO_F[f][2]];
                for (int64_t cc = 1; cc < r; ++cc) {
                    const double a = (double)(d - r), bb = (double)(r - cc), w = (double)cc;
                    project(a * A[0] + bb * B[0] + w * C[0], a * A[1] + bb * B[1] + w * C[1],
                            a * A[2] + bb * B[2] + w * C[2], verts + 3 * g.index(f, r, cc));
                }
            }
        }, 8);
    }
    if (tris) {
        // Per face, rows r = 0..d-1: "up" triangles for c = 0..r and "down"
        // triangles for c = 0..r-1, both counter-clockwise like the face.
        parallel_for(20 * d, [&](int64_t b, int64_t e) {
          for (int64_t j = b; j < e; ++j) {
            const int f = (int)(j / d);
            const int64_t r = j % d;
            int64_t t = (int64_t)f * d * d + r * r;  // rows before r hold r^2 triangles
            {
                for (int64_t cc = 0; cc <= r; ++cc) {
                    int32_t *o = tris + 3 * t++;
                    o[0] = (int32_t)g.index(f, r, cc);
                    o[1] = (int32_t)g.index(f, r + 1, cc);
                    o[2] = (int32_t)g.index(f, r + 1, cc + 1);
                }
                for (int64_t cc = 0; cc < r; ++cc) {
                    int32_t *o = tris + 3 * t++;
                    o[0] = (int32_t)g.index(f, r, cc);
                    o[1] = (int32_t)g.index(f, r + 1, cc + 1);
                    o[2] = (int32_t)g.index(f, r, cc + 1);
                }
            }
          }
        }, 8);
    }
    return NX_OK;
}

// util.py:590-613: for each triangle (a,b,c) append b to row a, c to row b and
// a to row c, in triangle order.  Serial, as the reference, because the order
// of appends fixes each row's first entry.
int build_adjacency(const int32_t *tris, int64_t T, int64_t V, int32_t *adj)
{
    if (!tris || !adj || T < 0 || V < 0) return NX_EINVAL;
    std::fill(adj, adj + V * 6, -1);
    std::vector<uint8_t> fill((size_t)V, 0);
    for (int64_t t = 0; t < T; ++t) {
        const int32_t v[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
        for (int k = 0; k < 3; ++k) {
            const int32_t a = v[k], b = v[(k + 1) % 3];
            if (a < 0 || a >= V || b < 0 || b >= V) return NX_EINVAL;
            if (fill[a] >= 6) return NX_EINVAL;  // vertex of degree > 6
            adj[(int64_t)a * 6 + fill[a]++] = b;
        }
    }
    return NX_OK;
}

// util.py:638-662: walk each vertex's ring starting from its first entry;
// the next vertex is the first entry of the vertex's own (unsorted) row that
// is also a neighbour of the current one and is not the previous one
// (util.py:623-629).  Rows 0..11 have five neighbours, as the reference
// assumes.  Membership tests only need row sets, so reading the unsorted
// table is what a race-free run of the reference computes.
int sort_adjacency(int32_t *adj, int64_t V)
{
    if (!adj || V < 12) return NX_EINVAL;
    std::vector<int32_t> orig(adj, adj + V * 6);
    std::atomic<int> bad{0};
    parallel_for(V, [&](int64_t b0, int64_t e0) {
      for (int64_t idx = b0; idx < e0; ++idx) {
        const int deg = idx < 12 ? 5 : 6;
        const int32_t *row = orig.data() + idx * 6;
        int32_t visited[6] = {-1, -1, -1, -1, -1, -1};
        int32_t pv = (int32_t)idx, nv = row[0];
        for (int i = 0; i < deg - 1; ++i) {
            visited[i] = nv;
            int32_t next = -1;
            if (nv >= 0) {
                const int32_t *other = orig.data() + (int64_t)nv * 6;
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
            adj[idx * 6 + k] = visited[k];
            if (k < deg && visited[k] < 0) bad.store(1);
        }
      }
    });
    return bad.load() ? NX_EINVAL : NX_OK;
}

// Space-filling-curve order: 3D Morton key of the coordinates quantised to
// 21 bits per axis over [-1,1] (or the bounding box for other radii).
int sfc_order(const double *verts, int64_t V, int64_t *order)
{
    if (!verts || !order || V <= 0) return NX_EINVAL;
    double lo[3] = {verts[0], verts[1], verts[2]}, hi[3] = {verts[0], verts[1], verts[2]};
    for (int64_t i = 0; i < V; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], verts[3 * i + k]);
            hi[k] = std::max(hi[k], verts[3 * i + k]);
        }
    auto spread = [](uint64_t x) {
        x &= 0x1fffff;
        x = (x | x << 32) & 0x1f00000000ffffULL;
        x = (x | x << 16) & 0x1f0000ff0000ffULL;
        x = (x | x << 8) & 0x100f00f00f00f00fULL;
        x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
        x = (x | x << 2) & 0x1249249249249249ULL;
        return x;
    };
    std::vector<uint64_t> key((size_t)V);
    parallel_for(V, [&](int64_t b0, int64_t e0) {
      for (int64_t i = b0; i < e0; ++i) {
        uint64_t q[3];
        for (int k = 0; k < 3; ++k) {
            const double span = hi[k] - lo[k];
            double t = span > 0 ? (verts[3 * i + k] - lo[k]) / span : 0.0;
            q[k] = (uint64_t)std::min(2097151.0, std::max(0.0, t * 2097151.0));
        }
        key[i] = spread(q[0]) | spread(q[1]) << 1 | spread(q[2]) << 2;
      }
    });
    // Stable LSD radix sort of (key, index) on the 63 key bits, 9 passes of
    // 7 bits; ties keep the input order.
    std::vector<int64_t> tmp((size_t)V);
    std::vector<uint64_t> ktmp((size_t)V);
    std::iota(order, order + V, (int64_t)0);
    int64_t *src = order, *dst = tmp.data();
    uint64_t *ksrc = key.data(), *kdst = ktmp.data();
    for (int shift = 0; shift < 63; shift += 7) {
        int64_t cnt[129] = {0};
        for (int64_t i = 0; i < V; ++i) ++cnt[((ksrc[i] >> shift) & 127) + 1];
        for (int b = 0; b < 128; ++b) cnt[b + 1] += cnt[b];
        for (int64_t i = 0; i < V; ++i) {
            const int64_t pos = cnt[(ksrc[i] >> shift) & 127]++;
            dst[pos] = src[i];
            kdst[pos] = ksrc[i];
        }
        std::swap(src, dst);
        std::swap(ksrc, kdst);
    }
    if (src != order) std::copy(src, src + V, order);
    return NX_OK;
}

// Device neighbour layout.  order (nullable) maps new index -> old index;
// col is slot-major [6][V] in the new numbering and elen is [6][V]: each
// edge length computed as the reference's calc_distance (erosion.py:38-40)
// does, in FP64 (this file is built without FP contraction).
int ell_from_adjacency(const int32_t *adj, int64_t V, const double *verts,
                       const int64_t *order, int32_t *col, double *elen)
{
    if (!adj || !verts || !col || !elen || V <= 0 || V > INT32_MAX) return NX_EINVAL;
    std::vector<int32_t> inv;
    if (order) {
        inv.assign((size_t)V, -1);
        for (int64_t i = 0; i < V; ++i) {
            if (order[i] < 0 || order[i] >= V || inv[order[i]] >= 0) return NX_EINVAL;
            inv[order[i]] = (int32_t)i;
        }
    }
    std::atomic<int> bad{0};
    parallel_for(V, [&](int64_t b0, int64_t e0) {
        for (int64_t i = b0; i < e0; ++i) {
            const int64_t old = order ? order[i] : i;
            const double *p = verts + 3 * old;
            for (int k = 0; k < 6; ++k) {
                const int32_t n = adj[old * 6 + k];
                double *el = elen + (int64_t)k * V + i;
                if (n < -1 || n >= V) {
                    bad.store(1);
                    continue;
                }
                if (n == -1) {
                    col[(int64_t)k * V + i] = (int32_t)i;
                    el[0] = 0.0;
                } else {
                    const double *q = verts + 3 * (int64_t)n;
                    const double a = p[0] - q[0], b = p[1] - q[1], c = p[2] - q[2];
                    col[(int64_t)k * V + i] = order ? inv[n] : n;
                    el[0] = std::sqrt(a * a + b * b + c * c);
                }
            }
        }
    });
    return bad.load() ? NX_EINVAL : NX_OK;
}

}  // namespace nx
