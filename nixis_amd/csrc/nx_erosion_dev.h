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

#include "nx_ff.h"

namespace nx {

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

}  // namespace nx
