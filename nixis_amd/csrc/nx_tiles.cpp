// nx_tiles.cpp -- tile plan for erosion steps two per launch (host).
//
// A tile is a contiguous range of vertices in device order.  Its ring 1 is
// the set of neighbours outside the tile, its ring 2 the neighbours of ring 1
// outside tile and ring 1.  Two Jacobi steps of the tile need the fields of
// all three (erosion.py:197-279 reads only the one-ring of each vertex):
// step one on tile + ring 1, step two on the tile.  Tiles start at kTileMax
// vertices and are halved until tile + rings fit kLoadMax, so a
// space-filling-curve order (compact tiles, short perimeters) gives few,
// full tiles; a graph without locality fails with NX_ERANGE.
//
// Plan image (offsets 256-aligned, recorded in nx_tile_plan):
//   start  int32 [ntiles + 1]   first vertex of each tile (and V)
//   cs     int32 [ntiles]       tile + ring 1
//   load   int32 [ntiles]       tile + ring 1 + ring 2
//   rbase  int64 [ntiles + 1]   offset of each tile's rings in `ring`
//   ring   int32 [...]          global index of ring 1 then ring 2 vertices
//   cbase  int64 [ntiles + 1]   prefix sum of cs
//   lcol   uint16 [6 * sum cs]  per tile, slot-major [6][cs]: local index of
//                               each neighbour (tile: 0..T-1, ring 1: T.., ring 2 after)
//   lel    double [6 * sum cs]  per tile, slot-major: the slot's edge length
#include <algorithm>
#include <atomic>
#include <cstring>
#include <vector>

#include "nx_internal.h"
#include "nx_pair_dev.h"

namespace nx {

namespace {

struct Tile {
    int64_t a;
    int32_t T;
    std::vector<int32_t> r1, r2;
};

void sort_unique(std::vector<int32_t> &v)
{
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
}

// Rings of [a, a+T); false if they do not fit.
bool rings(const int32_t *col, int64_t V, int64_t a, int32_t T, Tile &t)
{
    const int64_t b = a + T;
    t.a = a;
    t.T = T;
    t.r1.clear();
    t.r2.clear();
    for (int k = 0; k < 6; ++k)
        for (int64_t v = a; v < b; ++v) {
            const int32_t n = col[(int64_t)k * V + v];
            if (n < a || n >= b) t.r1.push_back(n);
        }
    sort_unique(t.r1);
    if ((int64_t)T + (int64_t)t.r1.size() > kLoadMax) return false;
    for (int k = 0; k < 6; ++k)
        for (int32_t u : t.r1) {
            const int32_t n = col[(int64_t)k * V + u];
            if ((n < a || n >= b) && !std::binary_search(t.r1.begin(), t.r1.end(), n))
                t.r2.push_back(n);
        }
    sort_unique(t.r2);
    return (int64_t)T + (int64_t)t.r1.size() + (int64_t)t.r2.size() <= kLoadMax;
}

// Cut [a, a+len) into tiles that fit, halving as needed.
bool cut(const int32_t *col, int64_t V, int64_t a, int64_t len, std::vector<Tile> &out)
{
    Tile t;
    if (rings(col, V, a, (int32_t)len, t)) {
        out.push_back(std::move(t));
        return true;
    }
    if (len / 2 < kTileMin) return false;
    const int64_t h = len / 2;
    return cut(col, V, a, h, out) && cut(col, V, a + h, len - h, out);
}

inline int64_t a256(int64_t x) { return (x + 255) & ~(int64_t)255; }

}  // namespace

int tile_plan_build(const int32_t *col, const double *elen, int64_t V, void *out, int64_t cap,
                    nx_tile_plan *plan)
{
    if (!col || !elen || !plan || V <= 0 || V > INT32_MAX)
        return set_error(NX_EINVAL, "tile plan needs col, elen, plan and 0 < V < 2^31");
    for (int64_t i = 0; i < 6 * V; ++i)
        if (col[i] < 0 || col[i] >= V) return set_error(NX_EINVAL, "col[%lld] = %d outside 0..V-1", (long long)i, col[i]);

    // Chunks of kTileMax, each cut independently (in parallel), in order.
    const int64_t nchunks = (V + kTileMax - 1) / kTileMax;
    std::vector<std::vector<Tile>> parts((size_t)nchunks);
    std::vector<char> ok((size_t)nchunks, 1);
    parallel_for(nchunks, [&](int64_t c0, int64_t c1) {
        for (int64_t c = c0; c < c1; ++c) {
            const int64_t a = c * kTileMax;
            ok[(size_t)c] = cut(col, V, a, std::min<int64_t>(kTileMax, V - a), parts[(size_t)c]);
        }
    }, 16);
    for (int64_t c = 0; c < nchunks; ++c)
        if (!ok[(size_t)c])
            return set_error(NX_ERANGE, "vertices %lld.. do not cut into tiles whose two rings fit %d vertices",
                             (long long)(c * kTileMax), kLoadMax);
    std::vector<const Tile *> tiles;
    for (auto &p : parts)
        for (auto &t : p) tiles.push_back(&t);
    const int64_t nt = (int64_t)tiles.size();

    int64_t nring = 0, ncs = 0;
    int32_t max_tile = 0, max_load = 0;
    for (const Tile *t : tiles) {
        nring += (int64_t)(t->r1.size() + t->r2.size());
        ncs += t->T + (int64_t)t->r1.size();
        max_tile = std::max(max_tile, t->T);
        max_load = std::max<int32_t>(max_load, t->T + (int32_t)(t->r1.size() + t->r2.size()));
    }
    // Rings larger than the tiles themselves read more than two single
    // steps would: such a graph has no locality to exploit.
    if (nring > V)
        return set_error(NX_ERANGE, "tile rings hold %lld vertices for %lld owned: no locality",
                         (long long)nring, (long long)V);
    nx_tile_plan q;
    std::memset(&q, 0, sizeof(q));
    q.V = V;
    q.ntiles = nt;
    q.max_tile = max_tile;
    q.max_load = max_load;
    int64_t o = 0;
    q.off_start = o; o = a256(o + 4 * (nt + 1));
    q.off_cs = o;    o = a256(o + 4 * nt);
    q.off_load = o;  o = a256(o + 4 * nt);
    q.off_rbase = o; o = a256(o + 8 * (nt + 1));
    q.off_ring = o;  o = a256(o + 4 * nring);
    q.off_cbase = o; o = a256(o + 8 * (nt + 1));
    q.off_lcol = o;  o = a256(o + 2 * 6 * ncs);
    q.off_lel = o;   o = a256(o + 8 * 6 * ncs);
    q.bytes = o;
    q.blob = plan->blob;
    *plan = q;
    if (!out || cap < q.bytes) return NX_OK;

    char *img = (char *)out;
    auto *start = (int32_t *)(img + q.off_start);
    auto *cs = (int32_t *)(img + q.off_cs);
    auto *load = (int32_t *)(img + q.off_load);
    auto *rbase = (int64_t *)(img + q.off_rbase);
    auto *ring = (int32_t *)(img + q.off_ring);
    auto *cbase = (int64_t *)(img + q.off_cbase);
    auto *lcol = (uint16_t *)(img + q.off_lcol);
    auto *lel = (double *)(img + q.off_lel);
    rbase[0] = 0;
    cbase[0] = 0;
    for (int64_t i = 0; i < nt; ++i) {
        const Tile &t = *tiles[(size_t)i];
        start[i] = (int32_t)t.a;
        cs[i] = t.T + (int32_t)t.r1.size();
        load[i] = cs[i] + (int32_t)t.r2.size();
        rbase[i + 1] = rbase[i] + (int64_t)(t.r1.size() + t.r2.size());
        cbase[i + 1] = cbase[i] + cs[i];
    }
    start[nt] = (int32_t)V;
    std::atomic<int> bad{0};
    parallel_for(nt, [&](int64_t t0, int64_t t1) {
        for (int64_t i = t0; i < t1; ++i) {
            const Tile &t = *tiles[(size_t)i];
            const int64_t a = t.a, b = t.a + t.T;
            const int32_t r1n = (int32_t)t.r1.size();
            std::copy(t.r1.begin(), t.r1.end(), ring + rbase[i]);
            std::copy(t.r2.begin(), t.r2.end(), ring + rbase[i] + r1n);
            auto local = [&](int32_t n) -> int32_t {
                if (n >= a && n < b) return (int32_t)(n - a);
                auto it = std::lower_bound(t.r1.begin(), t.r1.end(), n);
                if (it != t.r1.end() && *it == n) return t.T + (int32_t)(it - t.r1.begin());
                it = std::lower_bound(t.r2.begin(), t.r2.end(), n);
                if (it != t.r2.end() && *it == n) return t.T + r1n + (int32_t)(it - t.r2.begin());
                return -1;
            };
            const int32_t c = cs[i];
            for (int32_t j = 0; j < c; ++j) {
                const int64_t v = j < t.T ? a + j : t.r1[(size_t)(j - t.T)];
                for (int k = 0; k < 6; ++k) {
                    const int64_t e = 6 * cbase[i] + (int64_t)k * c + j;
                    const int32_t l = local(col[(int64_t)k * V + v]);
                    if (l < 0) bad.store(1);
                    lcol[e] = (uint16_t)(l < 0 ? 0 : l);
                    lel[e] = elen[(int64_t)k * V + v];
                }
            }
        }
    }, 64);
    if (bad.load()) return set_error(NX_EINVAL, "tile plan: a neighbour fell outside its tile's rings");
    return NX_OK;
}

}  // namespace nx
