// nx_internal.h -- declarations shared by the library's translation units.
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/nixis.h"

namespace nx {

// Record a message for nx_last_error() and return `code`.
int set_error(int code, const char *fmt, ...);

// Static-partition parallel loop over [0, n) on host threads (the library
// does not link an OpenMP runtime).  fn(begin, end) handles one chunk.
template <class F>
void parallel_for(int64_t n, F fn, int64_t min_chunk = 4096)
{
    if (n <= 0) return;
    int64_t hw = (int64_t)std::max(1u, std::thread::hardware_concurrency());
    int64_t nt = std::min<int64_t>(hw, std::max<int64_t>(1, n / min_chunk));
    if (nt <= 1) {
        fn((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve((size_t)nt);
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t b = n * t / nt, e = n * (t + 1) / nt;
        th.emplace_back([=]() { fn(b, e); });
    }
    for (auto &x : th) x.join();
}

// Host-side mesh / neighbour-table helpers (nx_mesh.cpp).
int icosphere_counts(int64_t d, int64_t *V, int64_t *T);
int icosphere(int64_t d, double *verts, int32_t *tris);
int build_adjacency(const int32_t *tris, int64_t T, int64_t V, int32_t *adj);
int sort_adjacency(int32_t *adj, int64_t V);
int sfc_order(const double *verts, int64_t V, int64_t *order);
int ell_from_adjacency(const int32_t *adj, int64_t V, const double *verts,
                       const int64_t *order, int32_t *col, double *elen);

// Tile plan for two steps per launch (nx_tiles.cpp).
int tile_plan_build(const int32_t *col, const double *elen, int64_t V, int K, void *out,
                    int64_t cap, nx_tile_plan *plan, int chunk = 0);

// Host-side permutation tables (nx_api.cpp).
int init_tables(int64_t seed, int32_t *perm, int32_t *pgi);

}  // namespace nx
