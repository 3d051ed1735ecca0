// nx_pair_dev.h -- two erosion steps over one tile of a tile plan.
//
// A tile owns vertices [start, start+T) in device order; ring 1 holds its
// outside neighbours and ring 2 theirs (nx_tiles.cpp builds the plan).  h and
// w of all three are staged in LDS, the first step runs on tile + ring 1
// from LDS, its h and w go back to LDS, and the second step runs on the tile
// and is written to the other buffer.  Ring-1 vertices are computed by every
// tile that touches them, with the same inputs and arithmetic
// (nx_erosion_dev.h), so the values agree and the result equals two
// one-step launches bit for bit.
#pragma once

#include <stdint.h>

#include "../../include/nixis.h"
#include "nx_erosion_dev.h"

namespace nx {

constexpr int kTileMax = 1024;      // owned vertices per tile
constexpr int kLoadMax = 2048;      // tile + both rings (two steps per launch)
// Two-step plans start from tiles of kPairChunk vertices and keep tile + ring
// 1 (the vertices of the first step) within kStepOneMax, the pair kernel's
// block size, so every first-step vertex has its own thread and its
// neighbour slots and lengths load before the block's first barrier.
constexpr int kPairChunk = 768;
constexpr int kStepOneMax = 1024;
constexpr int kTileMin = 32;        // smallest tile the planner cuts to
constexpr int kMultiSteps = 4;      // steps per launch of the multi-step kernel
constexpr int kLoadMaxMulti = 3072; // tile + its kMultiSteps rings
// Four-step plans keep the first step's vertices (tile + rings 1..3) within
// kQuadStepOne, two per thread of a 1024-thread block, and tile + rings
// 1..4 within kQuadLoadMax, so that both LDS buffers of erode_tile_quad fit
// two blocks per CU; they start from tiles of kQuadChunk vertices.
constexpr int kQuadStepOne = 2048;
constexpr int kQuadLoadMax = 2688;
constexpr int kQuadChunk = 1152;

// Device view of a plan image (pointers into nx_tile_plan::blob).
struct TilePlanDev {
    const int32_t *start;    // [ntiles + 1]
    const int32_t *cs;       // tile + ring 1 per tile
    const int32_t *load;     // tile + ring 1 + ring 2 per tile
    const int64_t *rbase;    // [ntiles + 1] into ring
    const int32_t *ring;     // global indices: ring 1 then ring 2, per tile
    const int64_t *cbase;    // [ntiles + 1] prefix of cs
    const uint16_t *lcol;    // slot codes (slot_nbr, slot_len), [6][cs] per tile
    const double *lel;       // edge lengths, [columns][cs] per tile from ebase
    const int32_t *pref;     // [ntiles][steps + 1]: tile, tile + ring 1, ...
    const int64_t *ebase;    // [ntiles + 1] into lel
    int64_t ntiles;
    int32_t steps;
};

inline TilePlanDev plan_dev(const nx_tile_plan &q)
{
    const char *b = (const char *)q.blob;
    TilePlanDev P;
    P.start = (const int32_t *)(b + q.off_start);
    P.cs = (const int32_t *)(b + q.off_cs);
    P.load = (const int32_t *)(b + q.off_load);
    P.rbase = (const int64_t *)(b + q.off_rbase);
    P.ring = (const int32_t *)(b + q.off_ring);
    P.cbase = (const int64_t *)(b + q.off_cbase);
    P.lcol = (const uint16_t *)(b + q.off_lcol);
    P.lel = (const double *)(b + q.off_lel);
    P.pref = (const int32_t *)(b + q.off_pref);
    P.ebase = (const int64_t *)(b + q.off_ebase);
    P.ntiles = q.ntiles;
    P.steps = q.steps;
    return P;
}

// Slot codes of a tile plan (nx_tiles.cpp, tile_edges): bits 0-11 are the
// neighbour's local index, bits 12-14 the column of the edge's length and
// bit 15 tells whether vertex li itself (set) or the neighbour holds it, so
// each undirected edge is read from one place by both of its ends.  A
// tile's own vertices (local 0..T-1) hold at most kLenFull lengths: columns
// 0..kLenFull-1 of the length block have an entry for each of the cs
// first-step vertices, later columns only for the ring vertices T..cs-1.
constexpr int kLenFull = 3;
NX_EHD int mul_small(int a, int b)  // a, b < 2^23 (a column and a vertex count)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __mul24(a, b);
#else
    return a * b;
#endif
}
NX_EHD int len_index(int col, int owner, int cs, int T)
{
    return col < kLenFull ? mul_small(col, cs) + owner
                          : kLenFull * cs + mul_small(col - kLenFull, cs - T) + (owner - T);
}
NX_EHD int slot_nbr(int code) { return code & 4095; }
template <class P>
NX_EHD double slot_len(int code, int li, int cs, int T, P len)
{
    const int owner = (code & 0x8000) ? li : (code & 4095);
    return len[len_index((code >> 12) & 7, owner, cs, T)];
}

#if defined(__HIPCC__)
// Two steps of tile `tile`, reading buffer p and writing buffer p ^ 1.
// THREADS is the block size; KEEP holds the tile's neighbour indices and
// edge lengths in registers between the steps (otherwise the second step
// reads them again); the kept vertices' data is loaded before the first
// barrier, while the field gathers are in flight.  A is the block's LDS
// staging area; B, when given, receives step one's h and w (one barrier
// fewer than writing them back into A).
template <int THREADS, bool KEEP, int RULE = 3>
__device__ __forceinline__ void erode_tile_pair(double2 *A, double2 *B, double2 *hw0, double2 *hw1,
                                                double *s0, double *s1, int p,
                                                const TilePlanDev &P, int64_t tile,
                                                double rain_mid, double rain_next)
{
    constexpr int kLoadPer = (kLoadMax + THREADS - 1) / THREADS;
    constexpr int kOwnPer = (kTileMax + THREADS - 1) / THREADS;
    constexpr int kKeep = KEEP ? kOwnPer : 1;
    const auto hw_src = gmem(p ? hw1 : hw0);
    const auto s_src = gmem(p ? s1 : s0);
    const auto hw_dst = gmem(p ? hw0 : hw1);
    const auto s_dst = gmem(p ? s0 : s1);
    const int64_t a = gmem(P.start)[tile];
    const int T = (int)(gmem(P.start)[tile + 1] - a);
    const int cs = gmem(P.cs)[tile], L = gmem(P.load)[tile];
    const auto ring = gmem(P.ring) + gmem(P.rbase)[tile];
    const int64_t cb = gmem(P.cbase)[tile];
    const auto lcol = gmem(P.lcol) + 6 * cb;
    const auto lel = gmem(P.lel) + gmem(P.ebase)[tile];

    for (int li = threadIdx.x; li < L; li += THREADS)
        A[li] = hw_src[li < T ? a + li : ring[li - T]];
    double sed[kLoadPer];
#pragma unroll
    for (int m = 0; m < kLoadPer; ++m) {
        const int li = threadIdx.x + m * THREADS;
        sed[m] = li < cs ? s_src[li < T ? a + li : ring[li - T]] : 0.0;
    }
    // The kept vertices' neighbour slots and lengths are loaded now, while
    // the gathers above are in flight, rather than after the barrier.
    int nbr[kKeep][6];
    double len[kKeep][6];
    if (KEEP) {
#pragma unroll
        for (int m = 0; m < kKeep; ++m) {
            const int li = threadIdx.x + m * THREADS;
            if (li < cs && m < kOwnPer) {
                int code[6];
#pragma unroll
                for (int q = 0; q < 6; ++q) code[q] = lcol[q * cs + li];
                // the six code loads go out together, before any length's
                // address depends on one of them
                asm volatile("" ::: "memory");
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    nbr[m][q] = slot_nbr(code[q]);
                    len[m][q] = slot_len(code[q], li, cs, T, lel);
                }
            }
        }
    }
    __syncthreads();

    double2 res[kLoadPer];
#pragma unroll
    for (int m = 0; m < kLoadPer; ++m) {
        const int li = threadIdx.x + m * THREADS;
        res[m] = make_double2(0.0, 0.0);
        if (li < cs) {
            const double2 me = A[li];
            double S = sed[m], W = me.y;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const bool kept = KEEP && m < kOwnPer;
                const int code = kept ? 0 : lcol[q * cs + li];
                const int n = kept ? nbr[m < kKeep ? m : 0][q] : slot_nbr(code);
                const double d = kept ? len[m < kKeep ? m : 0][q] : slot_len(code, li, cs, T, lel);
                const double2 nb = A[n];
                erode_accumulate<RULE>(me.x, nb.x, nb.y, d, S, W);
            }
            double ho, wo, so;
            erode_finish<RULE>(me.x, me.y, sed[m], S, W, rain_mid, ho, wo, so);
            if (B)
                B[li] = make_double2(ho, wo);
            else
                res[m] = make_double2(ho, wo);
            sed[m] = so;
        }
    }
    if (!B) {
        __syncthreads();
#pragma unroll
        for (int m = 0; m < kLoadPer; ++m) {
            const int li = threadIdx.x + m * THREADS;
            if (li < cs) A[li] = res[m];
        }
    }
    __syncthreads();
    const double2 *A2 = B ? B : A;

#pragma unroll
    for (int m = 0; m < kOwnPer; ++m) {
        const int li = threadIdx.x + m * THREADS;
        if (li < T) {
            const double2 me = A2[li];
            double S = sed[m], W = me.y;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const int code = KEEP ? 0 : lcol[q * cs + li];
                const int n = KEEP ? nbr[m < kKeep ? m : 0][q] : slot_nbr(code);
                const double d = KEEP ? len[m < kKeep ? m : 0][q] : slot_len(code, li, cs, T, lel);
                const double2 nb = A2[n];
                erode_accumulate<RULE>(me.x, nb.x, nb.y, d, S, W);
            }
            double ho, wo, so;
            erode_finish<RULE>(me.x, me.y, sed[m], S, W, rain_next, ho, wo, so);
            hw_dst[a + li] = make_double2(ho, wo);
            s_dst[a + li] = so;
        }
    }
}

// K steps of tile `tile` (a plan built for K steps), reading buffer p and
// writing buffer p ^ 1.  h and w of the tile and its K rings are staged in
// LDS; step j (1..K) runs on the first pref[K - j] local vertices (the tile
// and rings 1..K-j) and leaves its h and w in LDS for step j + 1, whose
// vertices need only those; step K writes the tile.  A vertex of ring m is
// computed by every tile that holds it, with the same inputs and arithmetic,
// so the result equals K one-step launches bit for bit.  KEEP holds the
// neighbour indices and edge lengths read in step 1 in registers for the
// later steps (otherwise they are read again, from cache).
template <int THREADS, int K, bool KEEP, int RULE = 3>
__device__ __forceinline__ void erode_tile_multi(double2 *A, double2 *hw0, double2 *hw1,
                                                 double *s0, double *s1, int p,
                                                 const TilePlanDev &P, int64_t tile,
                                                 double rain_mid, double rain_next)
{
    constexpr int kPer = (kLoadMaxMulti + THREADS - 1) / THREADS;
    constexpr int kKeep = KEEP ? kPer : 1;
    const auto hw_src = gmem(p ? hw1 : hw0);
    const auto s_src = gmem(p ? s1 : s0);
    const auto hw_dst = gmem(p ? hw0 : hw1);
    const auto s_dst = gmem(p ? s0 : s1);
    const int64_t a = gmem(P.start)[tile];
    const int T = (int)(gmem(P.start)[tile + 1] - a);
    const auto pref = gmem(P.pref) + tile * (K + 1);
    const int cs = pref[K - 1], L = pref[K];
    const auto ring = gmem(P.ring) + gmem(P.rbase)[tile];
    const int64_t cb = gmem(P.cbase)[tile];
    const auto lcol = gmem(P.lcol) + 6 * cb;
    const auto lel = gmem(P.lel) + gmem(P.ebase)[tile];

    for (int li = threadIdx.x; li < L; li += THREADS)
        A[li] = hw_src[li < T ? a + li : ring[li - T]];
    double sed[kPer];
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
        const int li = threadIdx.x + m * THREADS;
        sed[m] = li < cs ? s_src[li < T ? a + li : ring[li - T]] : 0.0;
    }
    __syncthreads();

    int nbr[kKeep][6];
    double len[kKeep][6];
    double2 res[kPer];
#pragma unroll
    for (int j = 1; j <= K; ++j) {
        const int n = pref[K - j];           // vertices of step j
        const double rain = j < K ? rain_mid : rain_next;
#pragma unroll
        for (int m = 0; m < kPer; ++m) {
            const int li = threadIdx.x + m * THREADS;
            if (li < n) {
                const double2 me = A[li];
                double S = sed[m], W = me.y;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    int nb_i;
                    double d;
                    if (KEEP && j > 1) {
                        nb_i = nbr[KEEP ? m : 0][q];
                        d = len[KEEP ? m : 0][q];
                    } else {
                        const int code = lcol[q * cs + li];
                        nb_i = slot_nbr(code);
                        d = slot_len(code, li, cs, T, lel);
                        if (KEEP) {
                            nbr[KEEP ? m : 0][q] = nb_i;
                            len[KEEP ? m : 0][q] = d;
                        }
                    }
                    const double2 nb = A[nb_i];
                    erode_accumulate<RULE>(me.x, nb.x, nb.y, d, S, W);
                }
                double ho, wo, so;
                erode_finish<RULE>(me.x, me.y, sed[m], S, W, rain, ho, wo, so);
                sed[m] = so;
                if (j < K) {
                    res[m] = make_double2(ho, wo);
                } else {
                    hw_dst[a + li] = make_double2(ho, wo);
                    s_dst[a + li] = so;
                }
            }
        }
        if (j < K) {
            __syncthreads();
#pragma unroll
            for (int m = 0; m < kPer; ++m) {
                const int li = threadIdx.x + m * THREADS;
                if (li < n) A[li] = res[m];
            }
            __syncthreads();
        }
    }
}

// Four steps of tile `tile` (a four-step plan within kQuadStepOne and
// kQuadLoadMax) by a 1024-thread block, reading buffer p and writing p ^ 1.
// Each thread keeps up to two first-step vertices (li and li + 1024): their
// sediment, neighbour indices (two to a word) and edge lengths are loaded
// before the first barrier, while the field gathers are in flight, and stay
// in registers for all four steps.  h and w of the tile and its four rings
// are staged in A; step 1 (tile + rings 1..3) writes B, step 2 (tile + rings
// 1..2) A, step 3 (tile + ring 1) B, and step 4 writes the tile: one barrier
// between steps.  Every vertex is computed from the same inputs with the
// same arithmetic as by one-step launches, so the results are bit for bit
// theirs.  NS = 3 or 2 runs the last three or two of those steps (a run's
// tail of steps that does not fill a four-step launch): the first step then
// covers tile + rings 1..2 or tile + ring 1, and only the vertices it reads
// are staged.
template <int RULE = 3, int NS = 4>
__device__ __forceinline__ void erode_tile_quad(double2 *A, double2 *B, double2 *hw0, double2 *hw1,
                                                double *s0, double *s1, int p,
                                                const TilePlanDev &P, int64_t tile,
                                                double rain_mid, double rain_next)
{
    constexpr int TH = 1024;
    constexpr int kPer = kQuadStepOne / TH;
    const auto hw_src = gmem(p ? hw1 : hw0);
    const auto s_src = gmem(p ? s1 : s0);
    const auto hw_dst = gmem(p ? hw0 : hw1);
    const auto s_dst = gmem(p ? s0 : s1);
    const int64_t a = gmem(P.start)[tile];
    const int T = (int)(gmem(P.start)[tile + 1] - a);
    const auto pref = gmem(P.pref) + tile * (kMultiSteps + 1);
    static_assert(NS >= 2 && NS <= kMultiSteps, "2..4 steps");
    const int c3 = pref[1], c2 = pref[2], c1 = pref[3], L = pref[4];
    // first-step set and staged set of this launch; slot codes and lengths
    // keep the plan's layout over the c1 first-step vertices
    const int cf = NS == 4 ? c1 : NS == 3 ? c2 : c3;
    const int cl = NS == 4 ? L : NS == 3 ? c1 : c2;
    const auto ring = gmem(P.ring) + gmem(P.rbase)[tile];
    const int64_t cb = gmem(P.cbase)[tile];
    const auto lcol = gmem(P.lcol) + 6 * cb;
    const auto lel = gmem(P.lel) + gmem(P.ebase)[tile];

    for (int li = threadIdx.x; li < cl; li += TH)
        A[li] = hw_src[li < T ? a + li : ring[li - T]];
    double sed[kPer];
    int code[kPer][6];
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
        const int li = threadIdx.x + m * TH;
        sed[m] = 0.0;
        if (li < cf) {
            sed[m] = s_src[li < T ? a + li : ring[li - T]];
#pragma unroll
            for (int q = 0; q < 6; ++q) code[m][q] = lcol[q * c1 + li];
        }
    }
    // the code loads go out together, before any length's address depends on one
    asm volatile("" ::: "memory");
    uint32_t nbp[kPer][3];
    double len[kPer][6];
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
        const int li = threadIdx.x + m * TH;
        if (li < cf) {
#pragma unroll
            for (int q = 0; q < 6; ++q) len[m][q] = slot_len(code[m][q], li, c1, T, lel);
#pragma unroll
            for (int q = 0; q < 3; ++q)
                nbp[m][q] = (uint32_t)slot_nbr(code[m][2 * q]) |
                            (uint32_t)slot_nbr(code[m][2 * q + 1]) << 16;
        }
    }
    __syncthreads();

    // One step over local vertices 0..n-1 from `src`; `emit` takes the results.
    auto step = [&](const double2 *src, int n, double rain, auto &&emit) {
#pragma unroll
        for (int m = 0; m < kPer; ++m) {
            const int li = threadIdx.x + m * TH;
            if (li < n) {
                const double2 me = src[li];
                double S = sed[m], W = me.y;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const int nb_i = (int)((nbp[m][q >> 1] >> (16 * (q & 1))) & 0xffffu);
                    const double2 nb = src[nb_i];
                    erode_accumulate<RULE>(me.x, nb.x, nb.y, len[m][q], S, W);
                }
                double ho, wo, so;
                erode_finish<RULE>(me.x, me.y, sed[m], S, W, rain, ho, wo, so);
                sed[m] = so;
                emit(li, ho, wo, so);
            }
        }
    };
    auto to_a = [&](int li, double h, double w, double) { A[li] = make_double2(h, w); };
    auto to_b = [&](int li, double h, double w, double) { B[li] = make_double2(h, w); };
    auto out = [&](int li, double h, double w, double s) {
        hw_dst[a + li] = make_double2(h, w);
        s_dst[a + li] = s;
    };
    if (NS == 4) {
        step(A, c1, rain_mid, to_b);
        __syncthreads();
        step(B, c2, rain_mid, to_a);
        __syncthreads();
        step(A, c3, rain_mid, to_b);
        __syncthreads();
        step(B, T, rain_next, out);
    } else if (NS == 3) {
        step(A, c2, rain_mid, to_b);
        __syncthreads();
        step(B, c3, rain_mid, to_a);
        __syncthreads();
        step(A, T, rain_next, out);
    } else {
        step(A, c3, rain_mid, to_b);
        __syncthreads();
        step(B, T, rain_next, out);
    }
}
#endif

}  // namespace nx
