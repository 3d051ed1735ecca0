// nx_erosion_dev.h -- one vertex of the erosion Jacobi step in float-float.
//
// Restates erosion.py:209-267 (the body of erosion_iteration3's prange loop):
//   sed_amt = sed[i] + sum over neighbours n of +-solubility * wat[n]
//   wat_amt = wat[i] + sum over neighbours n of +-wat[n] * d(i, n)
// with + when h[n] > h[i], - when h[n] < h[i] and nothing on a tie, then
//   h' = h - sed_amt,  s' = s + sed_amt,  w' = w + (wat_amt - wat_amt * evap)
//   if s' > cap * w':  h' += s' - cap * w';  s' -= s' - cap * w'.
// The neighbour sums are accumulated as two float-float sums (sum of +-w[n]
// and of +-w[n]*d) and the solubility factor is applied once to the first;
// all fields and constants are float-float, so the result tracks the FP64
// reference to ~2^-46 relative per step (FP32 alone diverges once
// neighbouring heights differ by less than its 24-bit resolution).
#pragma once

#include <stdint.h>

#include "nx_ff.h"

namespace nx {

// Per-planet state at the head of an erosion workspace.
struct ErosionMeta {
    int64_t it;           // steps taken so far (selects buffers and slots)
    int32_t E[3];         // exponent after step it-1, rotating by step % 3
    uint32_t big[3];      // 1 if some output of that step exceeded 2^64
    uint64_t load_max;    // FP64 max |input| bits while loading
};

// erosion.py:205-207, rounded from the FP64 constants to float-float.
struct ErosionConsts {
    ff sol;       // 0.01 / 320
    ff keep;      // 1 - 0.1 / 320  (water kept after evaporation)
    ff cap;       // 0.2 / 320
};

NX_HD ErosionConsts erosion_consts()
{
    ErosionConsts c;
    c.sol = ff_from_double(0.01 / 320);
    c.keep = ff_from_double(1.0 - 0.1 / 320);
    c.cap = ff_from_double(0.2 / 320);
    return c;
}

// Accumulate one neighbour (height hn, water wn, edge length d) into the two
// signed sums.  A self-neighbour (pentagon pad slot) ties and adds nothing.
NX_HD void erode_accumulate(ff h, ff hn, ff wn, ff d, ff &sum_w, ff &sum_wd)
{
    const bool up = ff_gt(hn, h);
    const bool down = ff_lt(hn, h);
    const ff wd = ff_mul(wn, d);
    ff tw = up ? wn : ff_neg(wn);
    ff td = up ? wd : ff_neg(wd);
    if (!(up || down)) {
        tw = ff{0.0f, 0.0f};
        td = ff{0.0f, 0.0f};
    }
    sum_w = ff_add(sum_w, tw);
    sum_wd = ff_add(sum_wd, td);
}

// Finish the vertex: returns the new (h, w, s); `rain` (already scaled) is
// added to the new water for the next step.
NX_HD void erode_finish(const ErosionConsts &c, ff h, ff w, ff s, ff sum_w, ff sum_wd, ff rain,
                        ff &h_out, ff &w_out, ff &s_out)
{
    const ff sed_amt = ff_add(s, ff_mul(c.sol, sum_w));
    const ff wat_amt = ff_add(w, sum_wd);
    ff hn = ff_sub(h, sed_amt);
    ff sn = ff_add(s, sed_amt);
    const ff wn = ff_add(w, ff_mul(wat_amt, c.keep));
    const ff cw = ff_mul(c.cap, wn);
    if (ff_gt(sn, cw)) {
        const ff ex = ff_sub(sn, cw);
        hn = ff_add(hn, ex);
        sn = ff_sub(sn, ex);
    }
    h_out = hn;
    s_out = sn;
    w_out = ff_add(wn, rain);
}

#if defined(__HIPCC__)
// Outputs above this magnitude trigger a 2^-64 rescale of the next step.
constexpr float kBig = 18446744073709551616.0f;  // 2^64

// Raise the step's overflow-guard flag if any lane of the wave holds a value
// above 2^64: one ballot per wave and a coherent store of the constant 1 by
// the first lane (concurrent stores of the same value need no atomic RMW).
__device__ __forceinline__ void flag_big(bool over, uint32_t *flag)
{
    if (__any(over) && (threadIdx.x & 63) == 0)
        __hip_atomic_store(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
// buffers (scaled by `sc`), update, write the destination buffers; returns
// the largest output magnitude for the overflow guard.
__device__ __forceinline__ float erode_vertex_step(int64_t i, int64_t V, const float4 *hw_src,
                                                   const float2 *s_src, float4 *hw_dst,
                                                   float2 *s_dst, const int32_t *col,
                                                   const float2 *elen, float sc, ff rain)
{
    const ErosionConsts c = erosion_consts();
    const float4 me = hw_src[i];
    const float2 ms = s_src[i];
    const ff h = ff_scale(ff{me.x, me.y}, sc), w = ff_scale(ff{me.z, me.w}, sc);
    const ff s = ff_scale(ff{ms.x, ms.y}, sc);
    ff sum_w{0.0f, 0.0f}, sum_wd{0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int32_t n = col[(int64_t)k * V + i];
        const float2 d = elen[(int64_t)k * V + i];
        const float4 nb = hw_src[n];
        erode_accumulate(h, ff_scale(ff{nb.x, nb.y}, sc), ff_scale(ff{nb.z, nb.w}, sc),
                         ff{d.x, d.y}, sum_w, sum_wd);
    }
    ff ho, wo, so;
    erode_finish(c, h, w, s, sum_w, sum_wd, rain, ho, wo, so);
    hw_dst[i] = make_float4(ho.hi, ho.lo, wo.hi, wo.lo);
    s_dst[i] = make_float2(so.hi, so.lo);
    return fmaxf(fabsf(ho.hi), fmaxf(fabsf(wo.hi), fabsf(so.hi)));
}
#endif

}  // namespace nx
