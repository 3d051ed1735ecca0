// nx_ff.h -- float-float ("double-float") arithmetic on FP32 pairs.
//
// A value is hi + lo with |lo| <= ulp(hi)/2, about 48 significant bits.
// The building blocks are the error-free transformations two_sum (Knuth)
// and two_prod via fused multiply-add; the sums are the "sloppy" double-word
// addition (error bounded relative to |a| + |b|, which is the same bound
// the reference's FP64 additions have), products are the usual
// hi*hi + cross terms.  Everything is plain FP32 arithmetic; fmaf is the
// only fused operation and it is written explicitly (the build disables
// FP contraction so no other FMA is formed).
//
// Usable from host code (C++ compiler, for CPU checks of the exact kernel
// arithmetic) and device code.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define NX_HD __host__ __device__ __forceinline__
#else
#define NX_HD inline
#endif

namespace nx {

struct ff {
    float hi, lo;
};

NX_HD ff ff_make(float hi, float lo) { return ff{hi, lo}; }

// Nearest float-float to a double.
NX_HD ff ff_from_double(double x)
{
    const float hi = (float)x;
    return ff{hi, (float)(x - (double)hi)};
}

NX_HD double ff_to_double(ff a) { return (double)a.hi + (double)a.lo; }

NX_HD ff two_sum(float a, float b)
{
    const float s = a + b;
    const float bb = s - a;
    const float e = (a - (s - bb)) + (b - bb);
    return ff{s, e};
}

// Requires |a| >= |b| (or a == 0).
NX_HD ff fast_two_sum(float a, float b)
{
    const float s = a + b;
    const float e = b - (s - a);
    return ff{s, e};
}

NX_HD ff ff_add(ff a, ff b)
{
    const ff s = two_sum(a.hi, b.hi);
    const float e = s.lo + (a.lo + b.lo);
    return fast_two_sum(s.hi, e);
}

NX_HD ff ff_neg(ff a) { return ff{-a.hi, -a.lo}; }

NX_HD ff ff_sub(ff a, ff b) { return ff_add(a, ff_neg(b)); }

NX_HD ff ff_mul(ff a, ff b)
{
    const float p = a.hi * b.hi;
    float e = fmaf(a.hi, b.hi, -p);
    e = fmaf(a.hi, b.lo, fmaf(a.lo, b.hi, e));
    return fast_two_sum(p, e);
}

// Exact scaling by 2^k (k small enough not to overflow).
NX_HD ff ff_ldexp(ff a, int k) { return ff{ldexpf(a.hi, k), ldexpf(a.lo, k)}; }

NX_HD ff ff_scale(ff a, float pow2) { return ff{a.hi * pow2, a.lo * pow2}; }

// Comparisons of normalised pairs (lexicographic on hi, then lo).
NX_HD bool ff_gt(ff a, ff b) { return a.hi > b.hi || (a.hi == b.hi && a.lo > b.lo); }
NX_HD bool ff_lt(ff a, ff b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }

}  // namespace nx
