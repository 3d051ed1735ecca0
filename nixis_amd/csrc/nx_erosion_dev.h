// nx_erosion_dev.h -- one vertex of the erosion Jacobi step (erosion.py:209-267).
//
// The step is computed in FP64 in the reference's operation order, so every
// result is the reference's bit for bit:
//   S = sed[i], W = wat[i]
//   for each neighbour n (table order; a pad slot points at i itself, and a
//   tie adds nothing, exactly like the reference skipping -1):
//     slope = (h[n] - h[i]) / (d + 0.00001)                 (erosion.py:34-36)
//     slope > 0:  S += sol * wat[n];  W += wat[n] * d
//     slope < 0:  S -= sol * wat[n];  W -= wat[n] * d
//   h' = h - S;  s' = s + S;  w' = w + (W - W * evap)
//   if s' > cap * w':  h' += s' - cap * w';  s' -= s' - cap * w'
// The sign of the slope is the sign of h[n] - h[i] except where the quotient
// underflows to zero; that needs |h[n] - h[i]| below 2^-1000, where the
// division is done exactly as the reference does it.
//
// Why FP64 rather than FP32 fields: FP32 flips h[n] > h[i] comparisons and
// ends far from the reference (tests/test_erosion_precision.py); float-float
// pairs meet the bar but cost 8-11 FP32 instructions per operation, while on
// gfx950 one FP64 add, multiply or FMA takes about 1.7 times the issue time
// of an FP32 one.  The storage is 8 bytes per value either way.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define NX_EHD __host__ __device__ __forceinline__
#else
#define NX_EHD inline
#endif

namespace nx {

// Per-planet state at the head of an erosion workspace.
struct ErosionMeta {
    int64_t it;           // steps taken so far
    int32_t buf;          // which of the two buffers holds the current fields
    int32_t pad;
};

// erosion.py:205-207
constexpr double kEvap = 0.1 / 320;
constexpr double kSol = 0.01 / 320;
constexpr double kCap = 0.2 / 320;

// RULE 3 is erosion.py:197-279 (erosion_iteration3); RULE 4 is
// erosion.py:301-381 (erosion_iteration4), the same step with every
// contribution clamped by max(x, 0).  numba evaluates max(x, 0) as Python
// does: 0 when 0 > x, else x (so -0.0 and NaN pass through).
NX_EHD double max0(double x) { return 0.0 > x ? 0.0 : x; }

// One neighbour: hn, wn its height and water, d the edge length.
template <int RULE = 3>
NX_EHD void erode_accumulate(double h, double hn, double wn, double d, double &S, double &W)
{
    double slope = hn - h;
    if (__builtin_expect(fabs(slope) < 0x1p-1000, 0)) slope = slope / (d + 0.00001);
    // Both branches add or subtract the same two products: computed once
    // here, so lanes that take different branches share them.
    const double ts = RULE == 4 ? max0(kSol * wn) : kSol * wn;
    const double tw = RULE == 4 ? max0(wn * d) : wn * d;
    if (slope > 0) {
        S += ts;
        W += tw;
    } else if (slope < 0) {
        S -= ts;
        W -= tw;
    }
}

// Finish the vertex; `rain` (0 for none) is the next iteration's rainfall,
// added to the new water as the reference's water += rain would.
template <int RULE = 3>
NX_EHD void erode_finish(double h, double w, double s, double S, double W, double rain,
                         double &h_out, double &w_out, double &s_out)
{
    double hn, sn, wn;
    if (RULE == 4) {  // erosion.py:366-371
        hn = h - max0(S);
        sn = s + max0(S);
        wn = w + max0(W - W * kEvap);
        if (sn > (kCap * wn)) {
            hn += max0(sn - (kCap * wn));
            sn -= max0(sn - (kCap * wn));
        }
    } else {  // erosion.py:265-270
        hn = h - S;
        sn = s + S;
        wn = w + (W - W * kEvap);
        if (sn > (kCap * wn)) {
            hn += (sn - (kCap * wn));
            sn -= (sn - (kCap * wn));
        }
    }
    h_out = hn;
    s_out = sn;
    w_out = rain != 0.0 ? wn + rain : wn;
}

#if defined(__HIPCC__)
// Pointer into device global memory.  Pointers read from memory (planet
// descriptors) are generic to the compiler, which then emits flat accesses
// that it must also order against LDS; naming the address space gives plain
// global loads and stores.
template <class T>
__device__ __forceinline__ auto gmem(T *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (__attribute__((address_space(1))) T *)p;
#else
    return p;  // the host pass only parses device code
#endif
}

// Logical block index whose ranges are contiguous per XCD: the hardware
// hands consecutive workgroups to the 8 XCDs round-robin, so block b runs on
// XCD b % 8; giving each XCD one contiguous eighth of the vertices keeps the
// neighbour rows of a tile in that XCD's L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb)
{
    const int64_t per = (nb + 7) / 8;
    return (b % 8) * per + b / 8;
}

// One vertex of one step: gather the six neighbours from the source
// buffers, update, write the destination buffers.
template <int RULE = 3>
__device__ __forceinline__ void erode_vertex_step(int64_t i, int64_t V, const double2 *hw_src,
                                                  const double *s_src, double2 *hw_dst,
                                                  double *s_dst, const int32_t *col,
                                                  const double *elen, double rain)
{
    const double2 me = gmem(hw_src)[i];
    double S = gmem(s_src)[i], W = me.y;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int32_t n = gmem(col)[(int64_t)k * V + i];
        const double d = gmem(elen)[(int64_t)k * V + i];
        const double2 nb = gmem(hw_src)[n];
        erode_accumulate<RULE>(me.x, nb.x, nb.y, d, S, W);
    }
    double ho, wo, so;
    erode_finish<RULE>(me.x, me.y, gmem(s_src)[i], S, W, rain, ho, wo, so);
    gmem(hw_dst)[i] = make_double2(ho, wo);
    gmem(s_dst)[i] = so;
}
#endif

}  // namespace nx
