// nx_tiles.cpp -- tile plan for erosion steps K per launch (host).
//
// A tile is a contiguous range of vertices in device order.  Its ring 1 is
// the set of neighbours outside the tile, ring m + 1 the neighbours of ring m
// outside the tile and the earlier rings.  K Jacobi steps of the tile need
// the fields of the tile and K rings (erosion.py:197-279 reads only the
// one-ring of each vertex): step j runs on the tile and rings 1..K-j.  Tiles
// start at `chunk` vertices (by default kPairChunk for two steps,
// kQuadChunk for four; an autotuner may pick another per mesh size) and are
// halved until tile + rings fit the kernel's load limit and the first step's
// vertices (tile + rings 1..K-1) its step-one limit (kStepOneMax, or
// kQuadStepOne for four steps); a space-filling-curve order (compact tiles,
// short perimeters) gives few, full tiles; a graph without locality fails
// with NX_ERANGE.
//
// Plan image (offsets 256-aligned, recorded in nx_tile_plan):
//   start  int32 [ntiles + 1]   first vertex of each tile (and V)
//   cs     int32 [ntiles]       tile + rings 1..K-1 (the vertices of step 1)
//   load   int32 [ntiles]       tile + rings 1..K
//   rbase  int64 [ntiles + 1]   offset of each tile's rings in `ring`
//   ring   int32 [...]          global index of ring 1, ring 2, ... vertices
//   cbase  int64 [ntiles + 1]   prefix sum of cs
//   lcol   uint16 [6 * sum cs]  per tile, slot-major [6][cs]: the slot code
//                               (tile_edges): neighbour's local index (tile:
//                               0..T-1, then the rings), the column of its edge
//                               length and whether this vertex owns it
//   lel    double [...]         per tile, len_block(K_t, cs, T) doubles: each
//                               undirected edge length once, in its owner's
//                               column (tile_edges, len_index)
//   pref   int32 [ntiles][K+1]  tile, tile + ring 1, ..., tile + rings 1..K
//   ebase  int64 [ntiles + 1]   offset of each tile's lengths in `lel`
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
    std::vector<std::vector<int32_t>> r;  // rings 1..K, each sorted
    int32_t ring_total() const
    {
        int32_t n = 0;
        for (auto &x : r) n += (int32_t)x.size();
        return n;
    }
};

void sort_unique(std::vector<int32_t> &v)
{
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
}

// n in one of the finished (sorted) rings 1..m.
bool in_rings(const Tile &t, int m, int32_t n)
{
    for (int q = 0; q < m; ++q)
        if (std::binary_search(t.r[(size_t)q].begin(), t.r[(size_t)q].end(), n)) return true;
    return false;
}

// K rings of [a, a+T); false if tile + rings exceed `limit` or the first
// step's vertices (tile + rings 1..K-1) exceed `step1_limit`.
bool rings(const int32_t *col, int64_t V, int64_t a, int32_t T, int K, int limit, int step1_limit,
           Tile &t)
{
    const int64_t b = a + T;
    t.a = a;
    t.T = T;
    t.r.assign((size_t)K, {});
    int64_t total = T;
    for (int m = 0; m < K; ++m) {
        std::vector<int32_t> &cur = t.r[(size_t)m];
        auto add = [&](int32_t n) {
            if ((n < a || n >= b) && !in_rings(t, m, n)) cur.push_back(n);
        };
        if (m == 0) {
            for (int k = 0; k < 6; ++k)
                for (int64_t v = a; v < b; ++v) add(col[(int64_t)k * V + v]);
        } else {
            const std::vector<int32_t> &prev = t.r[(size_t)m - 1];
            for (int k = 0; k < 6; ++k)
                for (int32_t u : prev) add(col[(int64_t)k * V + u]);
        }
        sort_unique(cur);
        total += (int64_t)cur.size();
        if (total > limit || (m == K - 2 && total > step1_limit)) return false;
    }
    return true;
}

// Cut [a, a+len) into tiles that fit, halving as needed.
bool cut(const int32_t *col, int64_t V, int64_t a, int64_t len, int K, int limit,
         int step1_limit, std::vector<Tile> &out)
{
    Tile t;
    if (rings(col, V, a, (int32_t)len, K, limit, step1_limit, t)) {
        out.push_back(std::move(t));
        return true;
    }
    if (len / 2 < kTileMin) return false;
    const int64_t h = len / 2;
    return cut(col, V, a, h, K, limit, step1_limit, out) &&
           cut(col, V, a + h, len - h, K, limit, step1_limit, out);
}

// Doubles of a tile's length block: kt columns over c step-1 vertices, the
// columns from kLenFull on over the ring vertices T..c-1 only.
inline int64_t len_block(int kt, int c, int T)
{
    return (int64_t)std::min(kt, kLenFull) * c + (int64_t)std::max(kt - kLenFull, 0) * (c - T);
}

// Edge lengths of one tile, stored once per undirected edge.  Every length
// the tile's step-1 vertices (local 0..c-1, the tile's own 0..T-1 first)
// read belongs to one of them: an edge to a vertex outside that set to the
// vertex inside it, an edge between two of them to one of the two.  The
// owners are chosen so that each of the tile's own vertices holds at most
// kLenFull edges and the ring vertices as few as possible: greedy (an edge
// between an own and a ring vertex to the ring vertex, else to the end that
// holds fewer), then augmenting paths move ownership away from vertices over
// their cap (kLenFull for own vertices, kr for ring vertices, kr growing
// where no path exists).  Owner o's j-th owned edge (in slot order) has its
// length at len_index(j, o, c, T) (nx_pair_dev.h): columns below kLenFull
// hold all c vertices, later ones the ring vertices only.  The slot code is
// the neighbour's local index (bits 0-11), the length's column (bits 12-14)
// and bit 15 set when the slot's own vertex is the owner.  A pad slot (the
// vertex itself) is coded as owned, column 0: a tie adds nothing, whatever
// the length.  lc [6][c]: local neighbour index of each slot; len [6][c]:
// its length.  Returns the number of columns kt, or -1 when an own vertex
// cannot be held to kLenFull edges (a neighbour table that is not symmetric
// can force more on it); writes codes [6][c] and the len_block(kt, c, T)
// lengths when given.
int tile_edges(const int32_t *lc, const double *len, int c, int T, uint16_t *codes,
               double *lens)
{
    struct Edge {
        int32_t u, v;
        int8_t ku, kv;
        int32_t owner;
    };
    std::vector<Edge> E;
    std::vector<int32_t> own((size_t)c, 0);
    std::vector<std::vector<int32_t>> inc((size_t)c);   // internal edges at each vertex
    // a slot whose edge its own vertex must hold: the neighbour is outside
    // the step-1 set, or (a table that is not symmetric) has no slot back
    std::vector<int8_t> forced((size_t)6 * c, 0);
    auto back = [&](int32_t n, int32_t j) {
        for (int q = 0; q < 6; ++q)
            if (lc[(size_t)q * c + n] == j) return q;
        return -1;
    };
    for (int32_t j = 0; j < c; ++j)
        for (int k = 0; k < 6; ++k) {
            const int32_t n = lc[(size_t)k * c + j];
            if (n == j) continue;
            if (n >= c || back(n, j) < 0) {
                forced[(size_t)k * c + j] = 1;
                ++own[(size_t)j];
            }
        }
    for (int32_t j = 0; j < c; ++j)
        for (int k = 0; k < 6; ++k) {
            const int32_t n = lc[(size_t)k * c + j];
            if (n == j || forced[(size_t)k * c + j] || j > n) continue;
            const int kv = back(n, j);
            if (forced[(size_t)kv * c + n]) continue;
            const bool jr = j >= T, nr = n >= T;
            const int32_t o = jr != nr ? (jr ? j : n) : (own[(size_t)j] <= own[(size_t)n] ? j : n);
            ++own[(size_t)o];
            inc[(size_t)j].push_back((int32_t)E.size());
            inc[(size_t)n].push_back((int32_t)E.size());
            E.push_back(Edge{j, n, (int8_t)k, (int8_t)kv, o});
        }
    int kr = kLenFull;
    auto cap = [&](int32_t v) { return v < T ? kLenFull : kr; };
    std::vector<int32_t> prev((size_t)c);
    std::vector<int32_t> queue;
    for (int32_t s = 0; s < c; ++s) {
        while (own[(size_t)s] > cap(s)) {
            // breadth-first search over owned edges (x owns e: e can move to the
            // other end) for a vertex holding fewer than its cap
            std::fill(prev.begin(), prev.end(), -2);
            prev[(size_t)s] = -1;
            queue.assign(1, s);
            int32_t found = -1;
            for (size_t h = 0; h < queue.size() && found < 0; ++h) {
                const int32_t x = queue[h];
                for (int32_t e : inc[(size_t)x]) {
                    if (E[(size_t)e].owner != x) continue;
                    const int32_t y = E[(size_t)e].u == x ? E[(size_t)e].v : E[(size_t)e].u;
                    if (prev[(size_t)y] != -2) continue;
                    prev[(size_t)y] = e;
                    if (own[(size_t)y] < cap(y)) {
                        found = y;
                        break;
                    }
                    queue.push_back(y);
                }
            }
            if (found < 0) {
                if (s < T) return -1;
                ++kr;
                continue;
            }
            for (int32_t y = found; y != s;) {
                Edge &e = E[(size_t)prev[(size_t)y]];
                const int32_t x = e.u == y ? e.v : e.u;
                e.owner = y;
                y = x;
            }
            --own[(size_t)s];
            ++own[(size_t)found];
        }
    }
    int kt = 1;
    for (int32_t j = 0; j < c; ++j) kt = std::max(kt, own[(size_t)j]);
    if (!codes) return kt;
    // columns in slot order of each owner
    std::vector<int8_t> owns(forced);                  // slot (k, j) holds j's own edge
    for (const Edge &e : E) {
        if (e.owner == e.u)
            owns[(size_t)e.ku * c + e.u] = 1;
        else
            owns[(size_t)e.kv * c + e.v] = 1;
    }
    std::vector<int8_t> colof((size_t)6 * c, 0);
    std::fill(lens, lens + len_block(kt, c, T), 0.0);
    for (int32_t j = 0; j < c; ++j) {
        int m = 0;
        for (int k = 0; k < 6; ++k)
            if (owns[(size_t)k * c + j]) {
                colof[(size_t)k * c + j] = (int8_t)m;
                lens[len_index(m, j, c, T)] = len[(size_t)k * c + j];
                ++m;
            }
    }
    for (int32_t j = 0; j < c; ++j)
        for (int k = 0; k < 6; ++k) {
            const size_t sl = (size_t)k * c + j;
            const int32_t n = lc[sl];
            uint32_t code;
            if (n == j) {
                code = (uint32_t)j | 0x8000u;
            } else if (owns[sl]) {
                code = (uint32_t)n | (uint32_t)colof[sl] << 12 | 0x8000u;
            } else {
                const int q = back(n, j);   // n owns the edge, in its slot back to j
                code = (uint32_t)n | (uint32_t)colof[(size_t)q * c + n] << 12;
            }
            codes[sl] = (uint16_t)code;
        }
    return kt;
}

inline int64_t a256(int64_t x) { return (x + 255) & ~(int64_t)255; }

}  // namespace

int tile_plan_build(const int32_t *col, const double *elen, int64_t V, int K, void *out,
                    int64_t cap, nx_tile_plan *plan, int chunk_in)
{
    if (!col || !elen || !plan || V <= 0 || V > INT32_MAX)
        return set_error(NX_EINVAL, "tile plan needs col, elen, plan and 0 < V < 2^31");
    if (K != 2 && K != kMultiSteps)
        return set_error(NX_EINVAL, "tile plan for %d steps per launch: 2 or %d", K, kMultiSteps);
    // four-step kernels take any tile whose first-step set fits kQuadStepOne
    const int tile_max = K == 2 ? kTileMax : kQuadStepOne;
    if (chunk_in != 0 && (chunk_in < kTileMin || chunk_in > tile_max))
        return set_error(NX_EINVAL, "tile chunk %d outside %d..%d (0 for the default)", chunk_in,
                         kTileMin, tile_max);
    const int limit = K == 2 ? kLoadMax : kQuadLoadMax;
    const int step1_limit = K == 2 ? kStepOneMax : kQuadStepOne;
    const int64_t chunk = chunk_in ? chunk_in : (K == 2 ? kPairChunk : kQuadChunk);
    for (int64_t i = 0; i < 6 * V; ++i)
        if (col[i] < 0 || col[i] >= V) return set_error(NX_EINVAL, "col[%lld] = %d outside 0..V-1", (long long)i, col[i]);

    // Chunks of `chunk` vertices, each cut independently (in parallel), in order.
    const int64_t nchunks = (V + chunk - 1) / chunk;
    std::vector<std::vector<Tile>> parts((size_t)nchunks);
    std::vector<char> ok((size_t)nchunks, 1);
    parallel_for(nchunks, [&](int64_t c0, int64_t c1) {
        for (int64_t c = c0; c < c1; ++c) {
            const int64_t a = c * chunk;
            ok[(size_t)c] = cut(col, V, a, std::min<int64_t>(chunk, V - a), K, limit,
                                step1_limit, parts[(size_t)c]);
        }
    }, 16);
    for (int64_t c = 0; c < nchunks; ++c)
        if (!ok[(size_t)c])
            return set_error(NX_ERANGE,
                             "vertices %lld.. do not cut into tiles whose %d rings fit %d vertices "
                             "(tile + ring 1 within %d)",
                             (long long)(c * chunk), K, limit, step1_limit);
    std::vector<const Tile *> tiles;
    for (auto &p : parts)
        for (auto &t : p) tiles.push_back(&t);
    const int64_t nt = (int64_t)tiles.size();

    int64_t nring = 0, ncs = 0;
    int32_t max_tile = 0, max_load = 0, max_step1 = 0;
    for (const Tile *t : tiles) {
        nring += t->ring_total();
        ncs += t->T + t->ring_total() - (int64_t)t->r.back().size();
        max_tile = std::max(max_tile, t->T);
        max_load = std::max<int32_t>(max_load, t->T + t->ring_total());
        max_step1 = std::max<int32_t>(
            max_step1, (int32_t)(t->T + t->ring_total() - (int64_t)t->r.back().size()));
    }
    // Rings larger than the tiles themselves read more than K single steps
    // would: such a graph has no locality to exploit.
    if (nring > (int64_t)(K / 2) * V)
        return set_error(NX_ERANGE, "tile rings hold %lld vertices for %lld owned: no locality",
                         (long long)nring, (long long)V);

    // Local neighbour table [6][c] and the slots' lengths of tile i's step-1
    // vertices (the tile, then rings 1..K-1).
    std::atomic<int> bad{0};
    auto local_table = [&](const Tile &t, std::vector<int32_t> &flat, std::vector<int32_t> &lc,
                           std::vector<double> &ln) {
        const int64_t a = t.a, b = t.a + t.T;
        flat.clear();
        for (auto &x : t.r) flat.insert(flat.end(), x.begin(), x.end());
        auto local = [&](int32_t n) -> int32_t {
            if (n >= a && n < b) return (int32_t)(n - a);
            int32_t base = t.T;
            for (auto &x : t.r) {
                auto it = std::lower_bound(x.begin(), x.end(), n);
                if (it != x.end() && *it == n) return base + (int32_t)(it - x.begin());
                base += (int32_t)x.size();
            }
            return -1;
        };
        const int32_t c = t.T + t.ring_total() - (int32_t)t.r.back().size();
        lc.resize((size_t)6 * c);
        ln.resize((size_t)6 * c);
        for (int32_t j = 0; j < c; ++j) {
            const int64_t v = j < t.T ? a + j : flat[(size_t)(j - t.T)];
            for (int k = 0; k < 6; ++k) {
                const int32_t l = local(col[(int64_t)k * V + v]);
                if (l < 0) bad.store(1);
                lc[(size_t)k * c + j] = l < 0 ? j : l;
                ln[(size_t)k * c + j] = elen[(int64_t)k * V + v];
            }
        }
        return c;
    };
    // Sizing pass: each tile's length columns (tile_edges is deterministic,
    // so the writing pass below recomputes the same orientation).
    std::vector<int32_t> kts((size_t)nt, 0);
    parallel_for(nt, [&](int64_t t0, int64_t t1) {
        std::vector<int32_t> flat, lc;
        std::vector<double> ln;
        for (int64_t i = t0; i < t1; ++i) {
            const int32_t c = local_table(*tiles[(size_t)i], flat, lc, ln);
            kts[(size_t)i] = tile_edges(lc.data(), ln.data(), c, tiles[(size_t)i]->T, nullptr,
                                        nullptr);
        }
    }, 64);
    if (bad.load()) return set_error(NX_EINVAL, "tile plan: a neighbour fell outside its tile's rings");
    int64_t nlen = 0;
    for (int64_t i = 0; i < nt; ++i) {
        const Tile &t = *tiles[(size_t)i];
        if (kts[(size_t)i] < 0)
            return set_error(NX_ERANGE,
                             "tile plan: a tile's own vertex must hold more than %d edge lengths "
                             "(a neighbour table that is not symmetric)", kLenFull);
        nlen += len_block(kts[(size_t)i],
                          (int)(t.T + t.ring_total() - (int64_t)t.r.back().size()), t.T);
    }

    nx_tile_plan q;
    std::memset(&q, 0, sizeof(q));
    q.V = V;
    q.ntiles = nt;
    q.max_tile = max_tile;
    q.max_load = max_load;
    q.max_step1 = max_step1;
    int64_t o = 0;
    q.off_start = o; o = a256(o + 4 * (nt + 1));
    q.off_cs = o;    o = a256(o + 4 * nt);
    q.off_load = o;  o = a256(o + 4 * nt);
    q.off_rbase = o; o = a256(o + 8 * (nt + 1));
    q.off_ring = o;  o = a256(o + 4 * nring);
    q.off_cbase = o; o = a256(o + 8 * (nt + 1));
    q.off_lcol = o;  o = a256(o + 2 * 6 * ncs);
    q.off_lel = o;   o = a256(o + 8 * nlen);
    q.off_pref = o;  o = a256(o + 4 * (K + 1) * nt);
    q.off_ebase = o; o = a256(o + 8 * (nt + 1));
    q.steps = K;
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
    auto *pref = (int32_t *)(img + q.off_pref);
    auto *ebase = (int64_t *)(img + q.off_ebase);
    rbase[0] = 0;
    cbase[0] = 0;
    ebase[0] = 0;
    for (int64_t i = 0; i < nt; ++i) {
        const Tile &t = *tiles[(size_t)i];
        start[i] = (int32_t)t.a;
        int32_t n = t.T;
        pref[i * (K + 1)] = n;
        for (int m = 0; m < K; ++m) {
            n += (int32_t)t.r[(size_t)m].size();
            pref[i * (K + 1) + m + 1] = n;
        }
        load[i] = n;
        cs[i] = n - (int32_t)t.r.back().size();
        rbase[i + 1] = rbase[i] + t.ring_total();
        cbase[i + 1] = cbase[i] + cs[i];
        ebase[i + 1] = ebase[i] + len_block(kts[(size_t)i], cs[i], t.T);
    }
    start[nt] = (int32_t)V;
    parallel_for(nt, [&](int64_t t0, int64_t t1) {
        std::vector<int32_t> flat, lc;
        std::vector<double> ln;
        for (int64_t i = t0; i < t1; ++i) {
            const Tile &t = *tiles[(size_t)i];
            const int32_t c = local_table(t, flat, lc, ln);
            std::copy(flat.begin(), flat.end(), ring + rbase[i]);
            tile_edges(lc.data(), ln.data(), c, t.T, lcol + 6 * cbase[i], lel + ebase[i]);
        }
    }, 64);
    if (bad.load()) return set_error(NX_EINVAL, "tile plan: a neighbour fell outside its tile's rings");
    return NX_OK;
}

}  // namespace nx
