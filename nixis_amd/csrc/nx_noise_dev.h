// nx_noise_dev.h -- device code for 3D OpenSimplex noise (gfx950).
//
// Restates opensimplex.py:266-759 (noise3d) for one point per lane:
//   * skew, floor, lattice origin, in-cell coordinates, region and the two
//     extra vertices are computed in FP64 in the reference's operation order
//     (this file is compiled with -ffp-contract=off), so the selected lattice
//     points are bit-identical to the reference's;
//   * the lattice hash (opensimplex.py:125) reads byte tables in LDS;
//   * the per-point falloff and gradient dot product run in FP32.
//
// Layout of the evaluation (branch-free across the three regions, eight
// contribution slots, which is what the octahedron needs):
//   - six slots for unit-cube corners: 010, 001 (one bit set; out in region
//     1, the tetrahedron at 111), 101, 011 (two bits set; out in region 0,
//     the tetrahedron at 000), slot A = 100 or, in region 1, 111, and slot
//     B = 110 or, in region 0, 000;
//   - two slots hold the extra vertices chosen by the region logic, whose
//     lattice offsets and delta constants come from a 64-entry table indexed
//     by the region's comparison outcomes.
// The hash prefixes perm[x], perm[perm[x]+y] are shared between corners.
// The FP32 deltas are scaled by 1/sqrt(2), so attn / 2 is clamped to [0, 1]
// by the clamp modifier of the FMA that computes it (see contrib).
#pragma once

#include <stdint.h>

#include <cmath>
#include <cstring>

#include "nixis.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NX_ND __host__ __device__ __forceinline__
#else
#define NX_ND inline  // host build of the same code, for CPU checks
#endif

namespace nx {

// Extra vertices.  The region logic of opensimplex.py:302-686 ends in one of
// seven cases and the bit masks of the closest points, which fix the two
// extra lattice offsets (each component in -1..2, packed as offset+1 in two
// bits: ext0 in bits 0-5, ext1 in bits 6-11) and their FP32 delta constants
// o + (ox + oy + oz) / 3:
//   case 0: tetrahedron at 000, (0,0,0) among the closest two (mask = c)
//   case 1: tetrahedron at 000, otherwise (mask = a | b)
//   case 2: tetrahedron at 111, (1,1,1) among the closest two (mask = c)
//   case 3: tetrahedron at 111, otherwise (mask = a & b)
//   case 4: octahedron, both closest on the (1,1,1) side (mask = a & b)
//   case 5: octahedron, both on the (0,0,0) side (mask = a | b)
//   case 6: octahedron, one on each side (mask = far * 8 + near)
// An entry is 64 bytes: the kernels read the first three 16-byte words
// (delta constants and lattice offsets, each offset a whole word so it adds
// to a coordinate without being unpacked); the packed code is read only for
// selection records.
struct alignas(16) ExtEntry {
    float c0x, c0y, c0z;
    int32_t ox0;
    float c1x, c1y, c1z;
    int32_t ox1;
    int32_t oy0, oz0, oy1, oz1;
    int32_t code;
    int32_t pad[3];
};
// The case and mask are functions of five comparison outcomes per region,
// so the table is indexed by those bits (6-bit index):
//   bit 5 clear (tetrahedra, opensimplex.py:302-317, 417-434):
//     bit 0 = first score replacement (z replaces b), bit 1 = second (z
//     replaces a), bit 2 = a lattice point is closer than the two scores
//     ("near"), bit 3 = b's score beats a's (b > a at 000, b < a at 111),
//     bit 4 = tetrahedron at 111;
//   bit 5 set (octahedron, opensimplex.py:537-575): bit 0 = x + y > 1,
//     bit 1 = x + z > 1, bit 2 = y + z > 1, bit 3 = y + z replaces a, bit 4
//     = y + z replaces b.
constexpr int kExtIndex = 64;

struct ExtLut {
    ExtEntry e[kExtIndex];
};

constexpr int first_axis(int m) { return (m & 1) ? 0 : ((m & 2) ? 1 : 2); }
constexpr int first_clear(int m) { return !(m & 1) ? 0 : (!(m & 2) ? 1 : 2); }

constexpr ExtEntry ext_entry(int k, int m)
{
    int e0[3] = {0, 0, 0}, e1[3] = {0, 0, 0};
    const int cx = m & 1, cy = (m >> 1) & 1, cz = (m >> 2) & 1;
    switch (k) {
    case 0:  // single-bit c (opensimplex.py:317-351)
        e0[0] = cx ? 1 : -1;
        e1[0] = cx ? 1 : 0;
        e0[1] = cy ? 1 : (cx ? -1 : 0);
        e1[1] = cy ? 1 : (cx ? 0 : -1);
        e0[2] = cz ? 1 : 0;
        e1[2] = cz ? 1 : -1;
        break;
    case 1:  // two-bit c (opensimplex.py:353-384)
        for (int a = 0; a < 3; ++a) {
            e0[a] = (m >> a) & 1;
            e1[a] = ((m >> a) & 1) ? 1 : -1;
        }
        break;
    case 2:  // two-bit c (opensimplex.py:434-468)
        e0[0] = cx ? 2 : 0;
        e1[0] = cx ? 1 : 0;
        e0[1] = cy ? (cx ? 1 : 2) : 0;
        e1[1] = cy ? (cx ? 2 : 1) : 0;
        e0[2] = cz ? 1 : 0;
        e1[2] = cz ? 2 : 0;
        break;
    case 3:  // single-bit c (opensimplex.py:470-501)
        for (int a = 0; a < 3; ++a) {
            e0[a] = (m >> a) & 1;
            e1[a] = 2 * ((m >> a) & 1);
        }
        break;
    case 4:  // (1,1,1) and twice the shared axis (opensimplex.py:579-610)
        e0[0] = e0[1] = e0[2] = 1;
        e1[first_axis(m)] = 2;
        break;
    case 5:  // (0,0,0) and -1 on the omitted axis (opensimplex.py:611-641)
        for (int a = 0; a < 3; ++a) e1[a] = 1;
        e1[first_clear(m)] = -1;
        break;
    case 6: {  // (opensimplex.py:642-686)
        const int far = (m >> 3) & 7, near = m & 7;
        for (int a = 0; a < 3; ++a) e0[a] = 1;
        e0[first_clear(far)] = -1;
        e1[first_axis(near)] = 2;
        break;
    }
    default:
        break;
    }
    ExtEntry r{};
    r.code = (e0[0] + 1) | (e0[1] + 1) << 2 | (e0[2] + 1) << 4 | (e1[0] + 1) << 6 |
             (e1[1] + 1) << 8 | (e1[2] + 1) << 10;
    r.ox0 = e0[0];
    r.oy0 = e0[1];
    r.oz0 = e0[2];
    r.ox1 = e1[0];
    r.oy1 = e1[1];
    r.oz1 = e1[2];
    // o + (ox + oy + oz) / 3, scaled by 1 / sqrt(2) like the deltas
    const double R = 0.70710678118654752;
    const double t0 = (e0[0] + e0[1] + e0[2]) / 3.0;
    const double t1 = (e1[0] + e1[1] + e1[2]) / 3.0;
    r.c0x = (float)((e0[0] + t0) * R);
    r.c0y = (float)((e0[1] + t0) * R);
    r.c0z = (float)((e0[2] + t0) * R);
    r.c1x = (float)((e1[0] + t1) * R);
    r.c1y = (float)((e1[1] + t1) * R);
    r.c1z = (float)((e1[2] + t1) * R);
    return r;
}

// Case and mask of a table index (opensimplex.py:302-686 restated over the
// comparison bits; combinations the comparisons cannot produce, such as both
// replacements, map to some valid entry).
constexpr ExtEntry ext_of_index(int i)
{
    const bool b0 = i & 1, b1 = (i >> 1) & 1, b2 = (i >> 2) & 1, b3 = (i >> 3) & 1,
               b4 = (i >> 4) & 1;
    if (!((i >> 5) & 1)) {
        // a = (x, point 1) unless z replaced it, b = (y, point 2) unless z
        // replaced it; at 111 the points are the complements (6, 5, 3).
        const int ap = b1 && !b0 ? 4 : 1, bp = b0 ? 4 : 2;
        if (!b4) return b2 ? ext_entry(0, b3 ? bp : ap) : ext_entry(1, ap | bp);
        const int cap = 7 ^ ap, cbp = 7 ^ bp;
        return b2 ? ext_entry(2, b3 ? cbp : cap) : ext_entry(3, cap & cbp);
    }
    const bool f1 = b0, f2 = b1, f3 = b2, repA = b3, repB = b4 && !b3;
    const int ap = repA ? (f3 ? 6 : 1) : (f1 ? 3 : 4);
    const int bp = repB ? (f3 ? 6 : 1) : (f2 ? 5 : 2);
    const bool af = repA ? f3 : f1, bf = repB ? f3 : f2;
    if (af == bf) return af ? ext_entry(4, ap & bp) : ext_entry(5, ap | bp);
    return af ? ext_entry(6, ap * 8 + bp) : ext_entry(6, bp * 8 + ap);
}

constexpr ExtLut make_ext_lut()
{
    ExtLut t{};
    for (int i = 0; i < kExtIndex; ++i) t.e[i] = ext_of_index(i);
    return t;
}

// Tables in LDS.  The 24 gradient vectors come first, as 64-bit words (fp16
// gx | fp16 gy << 16 in the low half, the FP32 bits of gz in the high half),
// so a gradient's byte offset is its LDS address; the extra-vertex table
// follows (its fields fit the immediate offsets of the LDS loads), then perm
// and the gradient table as bytes with wrap-around padding: entry i + 1 holds
// table[i mod 256] for i in -1..(kWrap - 2), so every hash index used below
// (lattice coordinate mod 256 plus an offset in -1..2, plus a table value)
// lands in range without masking.  The gradient table holds 8 * (pgi / 3),
// the byte offset of the gradient's word, so a hash is an address.
constexpr int kWrap = 520;
struct alignas(16) NoiseTables {
    uint64_t g16[24];
    ExtEntry ext[kExtIndex];
    uint8_t perm[kWrap];
    uint8_t grad[kWrap];
};

// Seed-dependent part passed in kernel arguments.
struct NoiseTablesPacked {
    uint32_t perm[64];
    uint32_t grad[64];
    uint32_t g16[48];
};

// fp16 bits of the gradient components used by opensimplex.py:51-60.
inline uint32_t half_of_small(int v)
{
    const uint32_t sign = v < 0 ? 0x8000u : 0u;
    const int a = v < 0 ? -v : v;
    return sign | (a == 11 ? 0x4980u : 0x4400u);  // 11 = 1.375*2^3, 4 = 2^2
}

// Exact domain of the FP64 section: |skewed coordinate| < 2^30 keeps every
// lattice integer representable in int32 with margin.
constexpr double kNoiseCoordLimit = 1073741824.0;

NX_ND float half_to_float(uint32_t bits)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (float)__builtin_bit_cast(_Float16, (unsigned short)bits);
#else
    const uint32_t sign = (bits >> 15) & 1u, ex = (bits >> 10) & 31u, man = bits & 1023u;
    const float v = ex ? std::ldexp(1.0f + (float)man / 1024.0f, (int)ex - 15)
                       : std::ldexp((float)man / 1024.0f, -14);
    return sign ? -v : v;
#endif
}

#if defined(__HIPCC__)
__device__ const ExtLut kExtLutDev = make_ext_lut();

__device__ __forceinline__ void load_ext_table(NoiseTables &T)
{
    const float4 *src = reinterpret_cast<const float4 *>(kExtLutDev.e);
    float4 *dst = reinterpret_cast<float4 *>(T.ext);
    constexpr int n4 = (int)(sizeof(ExtLut) / sizeof(float4));
    for (int i = threadIdx.x; i < n4; i += blockDim.x) dst[i] = src[i];
}

__device__ __forceinline__ void load_tables(NoiseTables &T, const NoiseTablesPacked &P)
{
    const uint8_t *pb = reinterpret_cast<const uint8_t *>(P.perm);
    const uint8_t *gb = reinterpret_cast<const uint8_t *>(P.grad);
    uint32_t *hw = reinterpret_cast<uint32_t *>(T.g16);
    for (int i = threadIdx.x; i < kWrap; i += blockDim.x) {
        T.perm[i] = pb[(i + 255) & 255];
        T.grad[i] = gb[(i + 255) & 255];
    }
    for (int i = threadIdx.x; i < 48; i += blockDim.x) hw[i] = P.g16[i];
    load_ext_table(T);
}
#endif

// Clamp to [0, 1]; on the device this folds into the clamp modifier of the
// instruction that computes the operand.
NX_ND float clamp01(float a)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(a, 0.0f, 1.0f);
#else
    return std::fmin(std::fmax(a, 0.0f), 1.0f);
#endif
}

NX_ND float float_of_bits(uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(b);
#else
    float f;
    std::memcpy(&f, &b, 4);
    return f;
#endif
}

// One lattice contribution in FP32, accumulated into v.  The deltas are
// scaled by 1/sqrt(2), so attn / 2 = 1 - |d / sqrt(2)|^2 (attn <= 2 always),
// and clamping it to [0, 1] is the reference's attn > 0 test; `m` (1 for a
// selected point, 0 for one the region excludes) starts the chain, so an
// excluded point clamps to 0.  The slot adds (attn / 2)^4 * g.(d / sqrt(2)) =
// attn^4 (g.d) / kSumScale.  The fp16 gradient components feed the FMAs
// directly; `go` is a gradient table entry (the byte offset of the
// gradient's word).
NX_ND float contrib(float v, float dx, float dy, float dz, int go, const NoiseTables &T,
                    float m)
{
    float attn = fmaf(-dz, dz, fmaf(-dy, dy, fmaf(-dx, dx, m)));
    attn = clamp01(attn);
    const float a2 = attn * attn;
    const uint64_t w = *reinterpret_cast<const uint64_t *>(reinterpret_cast<const uint8_t *>(T.g16) + go);
    const uint32_t w0 = (uint32_t)w, w1 = (uint32_t)(w >> 32);
    const float dot = fmaf(half_to_float(w0 & 0xffffu), dx,
                           fmaf(half_to_float(w0 >> 16), dy, float_of_bits(w1) * dz));
    return fmaf(a2 * a2, dot, v);
}

// Gradient table entry (byte offset of the gradient's word) of lattice point
// (x, y, z) with x, y, z in -1..257 (a coordinate mod 256 plus an offset).
NX_ND int hash_full(const NoiseTables &T, int x, int y, int z)
{
    const int a = T.perm[1 + x];
    const int b = T.perm[1 + a + y];
    return T.grad[1 + b + z];
}

// noise3_sum's result times kSumScale = 16 sqrt(2) is the reference's sum of
// attn^4 (g . d) (the deltas are scaled by 1 / sqrt(2), attn by 1 / 2).
constexpr double kSumScaleD = 22.627416997969522;
constexpr float kSumScale = (float)kSumScaleD;

// Selection record for parity tests (same encoding as the oracle):
// sel[0..2] = xsb, ysb, zsb; sel[3] = region | ext0 offsets << 2 | ext1
// offsets << 8 (each offset + 1 in two bits per axis); sel[4] = checksum
// cs = cs * 31 + pgi of every selected lattice point in reference order.
struct NoiseSel {
    int v[5];
};

// Skewed coordinates of (x, y, z) (opensimplex.py:272-275, FP64, reference
// operation order).
struct Skewed {
    double x, y, z;
};

NX_ND Skewed skew3(double x, double y, double z)
{
    const double STRETCH = -1.0 / 6;
    const double so = (x + y + z) * STRETCH;
    return Skewed{x + so, y + so, z + so};
}

// Whether skewed coordinates lie in the exact domain of the FP64 section
// (false for NaN).  Outside it the noise value is meaningless; callers check
// the octave with the largest frequency, which bounds the others.
NX_ND bool in_domain(const Skewed &s)
{
    return (fabs(s.x) < kNoiseCoordLimit) & (fabs(s.y) < kNoiseCoordLimit) &
           (fabs(s.z) < kNoiseCoordLimit);
}

// Comparison outcomes as wave-wide lane masks, so the region logic below
// runs as scalar mask operations (one bit per lane) and only the final index
// bits are expanded per lane.  The host build keeps all-or-nothing masks.
typedef uint64_t LaneMask;
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ LaneMask lanes(bool b) { return __builtin_amdgcn_ballot_w64(b); }
__device__ __forceinline__ bool lane_bit(LaneMask m) { return __builtin_amdgcn_inverse_ballot_w64(m); }
// 2 v + (this lane's bit of m), as one add with m as the carry-in (a
// full-rate integer add, where a select per bit and the ORs joining the bits
// issue at about half rate).
__device__ __forceinline__ int shift_in(int v, LaneMask m)
{
    int r;
    LaneMask c;
    asm("v_addc_co_u32 %0, %1, %2, %2, %3" : "=v"(r), "=s"(c) : "v"(v), "s"(m));
    return r;
}
__device__ __forceinline__ int first_bit(LaneMask m)
{
    int r;
    LaneMask c;
    asm("v_addc_co_u32 %0, %1, 0, 0, %2" : "=v"(r), "=s"(c) : "s"(m));
    return r;
}
#else
inline LaneMask lanes(bool b) { return b ? ~0ull : 0ull; }
inline bool lane_bit(LaneMask m) { return m != 0; }
inline int shift_in(int v, LaneMask m) { return 2 * v + (m != 0 ? 1 : 0); }
inline int first_bit(LaneMask m) { return m != 0 ? 1 : 0; }
#endif

// Region of the in-cell coordinates (in_sum <= 1: tetrahedron at 000, >= 2:
// tetrahedron at 111, else octahedron) and the extra-vertex table index
// (ExtLut), from FP64 comparisons in the reference's form.  A NaN coordinate
// is outside the domain and flagged, so a complemented comparison stands for
// the opposite one.
struct RegionSel {
    int idx;
    bool inA, r1, octa;  // region 0, region 1, region 2
};

NX_ND RegionSel select_region(double xins, double yins, double zins, double in_sum)
{
    const LaneMask xy_ge = lanes(xins >= yins), xy_le = lanes(xins <= yins);
    const LaneMask zy_gt = lanes(zins > yins), zy_lt = lanes(zins < yins);
    const LaneMask zx_gt = lanes(zins > xins), zx_lt = lanes(zins < xins);
    // tetrahedron at (0,0,0): opensimplex.py:302-317.  a = (x, point 1),
    // b = (y, point 2); t1: z replaces b, t2: z replaces a.
    const double wA = 1 - in_sum;
    const LaneMask a1 = xy_ge & zy_gt;
    const LaneMask a2 = ~xy_ge & zx_gt;
    const LaneMask a0 = ~(a1 | a2);
    const LaneMask wx = lanes(wA > xins), wy = lanes(wA > yins), wz = lanes(wA > zins);
    const LaneMask nearA = (a1 & (wx | wz)) | (a2 & (wz | wy)) | (a0 & (wx | wy));
    // b's score beats a's: z > x (t1), y > z (t2), y > x
    const LaneMask gtA = (a1 & zx_gt) | (a2 & zy_lt) | (a0 & ~xy_ge);
    // tetrahedron at (1,1,1): opensimplex.py:417-434, comparisons reversed.
    const double wB = 3 - in_sum;
    const LaneMask b1 = xy_le & zy_lt;
    const LaneMask b2 = ~xy_le & zx_lt;
    const LaneMask b0 = ~(b1 | b2);
    const LaneMask vx = lanes(wB < xins), vy = lanes(wB < yins), vz = lanes(wB < zins);
    const LaneMask nearB = (b1 & (vx | vz)) | (b2 & (vz | vy)) | (b0 & (vx | vy));
    // b's score below a's: z < x (t1), y < z (t2), y < x
    const LaneMask ltB = (b1 & zx_lt) | (b2 & zy_gt) | (b0 & ~xy_le);
    // octahedron: opensimplex.py:537-575.  p - 1 when p > 1, else 1 - p
    // (1 - p equals -(p - 1) exactly in IEEE arithmetic, so this is |p - 1|).
    const double p1 = xins + yins, p2 = xins + zins, p3 = yins + zins;
    const LaneMask f1 = lanes(p1 > 1), f2 = lanes(p2 > 1), f3 = lanes(p3 > 1);
    const double asC = fabs(p1 - 1), bsC = fabs(p2 - 1), scC = fabs(p3 - 1);
    const LaneMask ab_le = lanes(asC <= bsC);
    const LaneMask repA = ab_le & lanes(asC < scC);
    const LaneMask repB = ~ab_le & lanes(bsC < scC);
    const LaneMask inA = lanes(in_sum <= 1), inB = lanes(in_sum >= 2);
    const LaneMask octa = ~(inA | inB);
    const LaneMask r1 = ~inA & inB;
    const LaneMask r0 = ~octa & ~r1;
    // table index bits (ExtLut)
    const LaneMask i0 = (octa & f1) | (r1 & b1) | (r0 & a1);
    const LaneMask i1 = (octa & f2) | (r1 & b2) | (r0 & a2);
    const LaneMask i2 = (octa & f3) | (r1 & nearB) | (r0 & nearA);
    const LaneMask i3 = (octa & repA) | (r1 & ltB) | (r0 & gtA);
    const LaneMask i4 = (octa & repB) | r1;
    RegionSel r;
    r.idx = shift_in(shift_in(shift_in(shift_in(shift_in(first_bit(octa), i4), i3), i2), i1), i0);
    r.inA = lane_bit(inA);
    r.r1 = lane_bit(r1);
    r.octa = lane_bit(octa);
    return r;
}

// Evaluate noise3d at the skewed coordinates (xs, ys, zs) of a point already
// multiplied by the octave's frequency.  Returns the FP32 sum of the
// contributions: noise3d's value is that divided by 103 (opensimplex.py:759).
// Every branch of the reference's region logic is evaluated with selects, so
// a wave never diverges.
template <bool SEL>
NX_ND float noise3_sum(double xs, double ys, double zs, const NoiseTables &T, NoiseSel *sel)
{
    // floor(xs) is the integer the reference's int(floor(xs)) holds, as a
    // double (exact inside the checked domain |xs| < 2^30), so it serves the
    // FP64 arithmetic directly and the int only the hashing.
    const double fx = floor(xs), fy = floor(ys), fz = floor(zs);
    const int xsb = (int)fx, ysb = (int)fy, zsb = (int)fz;
    const double xins = xs - fx, yins = ys - fy, zins = zs - fz;
    const double in_sum = xins + yins + zins;
    // The deltas to the cell origin, x - (xsb + sq) with sq = (xsb + ysb +
    // zsb) / 3 (opensimplex.py:283-299), only feed the FP32 section.  Since
    // x = xs - so and so = -(x + y + z) / 6, in_sum = -3 (so + sq), so the
    // delta is xins + in_sum / 3: the same value up to rounding, computed
    // below in FP32 from the FP64 in-cell coordinates.

    const RegionSel rs = select_region(xins, yins, zins, in_sum);
    const bool inA = rs.inA, r1 = rs.r1;
    const int region = inA ? 0 : (r1 ? 1 : 2);

    // ---- FP32 section -----------------------------------------------------
    constexpr float R = 0.70710678118654752f;   // 1 / sqrt(2)
    // every constant offset below is scaled by R as well (S is SQUISH * R)
    constexpr float S = (float)(1.0 / 3.0 * 0.70710678118654752);
    const float sr = (float)in_sum * S;
    const float d0x = fmaf((float)xins, R, sr), d0y = fmaf((float)yins, R, sr),
                d0z = fmaf((float)zins, R, sr);
    // (111's offset is 1 + 3 * SQUISH = 2 in every axis)
    constexpr float ONE = 0.70710678118654752f, TWO = (float)(2.0 * 0.70710678118654752);
    constexpr float ONE_S = (float)((1.0 + 1.0 / 3.0) * 0.70710678118654752);
    constexpr float ONE_2S = (float)((1.0 + 2.0 / 3.0) * 0.70710678118654752);
    constexpr float S2 = (float)(2.0 / 3.0 * 0.70710678118654752);
    const int xm = xsb & 255, ym = ysb & 255, zm = zsb & 255;

    // Shared hash prefixes for the unit cube (padded tables: index + 1).
    const int px0 = T.perm[1 + xm];
    const int px1 = T.perm[2 + xm];
    const int p00 = T.perm[1 + px0 + ym];
    const int p01 = T.perm[2 + px0 + ym];
    const int p10 = T.perm[1 + px1 + ym];
    const int p11 = T.perm[2 + px1 + ym];
    const int g100 = T.grad[1 + p10 + zm];
    const int g010 = T.grad[1 + p01 + zm];
    const int g001 = T.grad[2 + p00 + zm];
    const int g110 = T.grad[1 + p11 + zm];
    const int g101 = T.grad[2 + p10 + zm];
    const int g011 = T.grad[2 + p01 + zm];

    const float r1f = r1 ? 1.0f : 0.0f, inAf = inA ? 1.0f : 0.0f;
    const float oneBits = 1.0f - r1f;    // 010, 001
    const float twoBits = 1.0f - inAf;   // 101, 011

    // Corner c: delta = d0 - c - S * popcount(c).  Slot A holds 100, or 111
    // in region 1 (where the one-bit corners are out); slot B holds 110, or
    // 000 in region 0 (where the two-bit corners are out): eight slots cover
    // every region.  The two slots' deltas move from the first corner to the
    // second by an FMA with the region flag (0 or 1).
    const float ax = d0x - S, ay = d0y - S, az = d0z - S;
    const int gA = T.grad[1 + (r1 ? p11 + 1 : p10) + zm];
    float v = contrib(0.0f, fmaf(-r1f, TWO - ONE_S, ax - ONE), fmaf(-r1f, TWO - S, ay),
                      fmaf(-r1f, TWO - S, az), gA, T, 1.0f);
    v = contrib(v, ax, ay - ONE, az, g010, T, oneBits);
    v = contrib(v, ax, ay, az - ONE, g001, T, oneBits);
    const float bx = d0x - S2, by = d0y - S2, bz = d0z - S2;
    const int gB = T.grad[1 + (inA ? p00 : p11) + zm];
    v = contrib(v, fmaf(inAf, ONE_2S, bx - ONE), fmaf(inAf, ONE_2S, by - ONE),
                fmaf(inAf, S2, bz), gB, T, 1.0f);
    v = contrib(v, bx - ONE, by, bz - ONE, g101, T, twoBits);
    v = contrib(v, bx, by - ONE, bz - ONE, g011, T, twoBits);
    // Extra vertices from the table (three 16-byte LDS reads of the entry).
    const ExtEntry &E = T.ext[rs.idx];
    const int ge0 = hash_full(T, xm + E.ox0, ym + E.oy0, zm + E.oz0);
    const int ge1 = hash_full(T, xm + E.ox1, ym + E.oy1, zm + E.oz1);
    v = contrib(v, d0x - E.c0x, d0y - E.c0y, d0z - E.c0z, ge0, T, 1.0f);
    v = contrib(v, d0x - E.c1x, d0y - E.c1y, d0z - E.c1z, ge1, T, 1.0f);

    if (SEL) {
        // Reference evaluation order of the selected points; the checksum
        // accumulates pgi = 3 * gradient index = 3 * (table entry / 8).
        uint32_t cs = 0;
        auto add = [&cs](int g) { cs = cs * 31u + 3u * (uint32_t)(g >> 3); };
        if (region == 0) {
            add(T.grad[1 + p00 + zm]);
            add(g100);
            add(g010);
            add(g001);
        } else if (region == 1) {
            add(g110);
            add(g101);
            add(g011);
            add(T.grad[2 + p11 + zm]);
        } else {
            add(g100);
            add(g010);
            add(g001);
            add(g110);
            add(g101);
            add(g011);
        }
        add(ge0);
        add(ge1);
        sel->v[0] = xsb;
        sel->v[1] = ysb;
        sel->v[2] = zsb;
        const int code = E.code;
        sel->v[3] = region | (code & 63) << 2 | ((code >> 6) & 63) << 8;
        sel->v[4] = (int)cs;
    }
    return v;
}

// noise3d (opensimplex.py:266-759) at (x, y, z): the value in FP32; `bad` is
// set when the point is outside the exact domain.
template <bool SEL>
NX_ND float noise3(double x, double y, double z, const NoiseTables &T, bool &bad, NoiseSel *sel)
{
    const Skewed s = skew3(x, y, z);
    bad = bad | !in_domain(s);
    return noise3_sum<SEL>(s.x, s.y, s.z, T, sel) * (kSumScale / 103.0f);
}

// Per-octave scalars of one sample_octaves call (terrain.py:36-45): octave o
// adds ((noise + 1) * 0.5 * strength) * radius (terrain.py:28, 43), which is
// w_o + (w_o * kSumScale / 103) * sum_o with w_o = 0.5 * strength * radius
// and sum_o the FP32 sum of noise3_sum.  The w_o add up in FP64 (base); the
// weighted sums add up in FP32 over the octaves and join them at the end.
struct OctavePlan {
    double rough[NX_MAX_OCTAVES];  // frequency / radius of each octave
    float wsumf[NX_MAX_OCTAVES];   // w_o * kSumScale / 103 (FP32)
    double base;                   // sum of w_o over the octaves
    int32_t n_oct;
    int32_t check_oct;             // octave with the largest |rough|
    int32_t doubling;              // rough[o + 1] == 2 * rough[o] for every o
    int32_t pad;
};

// Whether doubling octave o's skewed coordinates gives octave o + 1's bit for
// bit, for a vertex (x, y, z) whose first octave has frequency r0, when the
// frequencies double exactly.  It does when none of the first octave's
// products rounds in the subnormal range, i.e. x * r0, y * r0, z * r0 and
// ((x + y) + z) * STRETCH are zero or at least 2^-1021 in magnitude: every
// later product is then the first's times a power of two with the same
// rounding, and sums scale exactly.  NaN fails the test.
NX_ND bool doubles_exactly(double x, double y, double z, double r0)
{
    const double lim = 0x1p-1021;
    const double X = x * r0, Y = y * r0, Z = z * r0;
    const double s = X + Y + Z;
    const double so = s * (-1.0 / 6);
    return ((x == 0) | (fabs(X) >= lim)) & ((y == 0) | (fabs(Y) >= lim)) &
           ((z == 0) | (fabs(Z) >= lim)) & ((s == 0) | (fabs(so) >= lim));
}

#if defined(__HIPCC__)
// The octave loop of terrain.py:40-45 for one vertex: returns acc plus every
// octave's weighted noise; `bad` is set when the octave with the largest
// frequency leaves the exact domain.  `on_octave(o, sum, sel)` sees each octave's
// contribution sum (the noise value times 103) and selection record.  The
// doubling shortcut is taken per wave (every lane doubles or none does), so
// the branch is uniform.
template <bool SEL, typename F>
__device__ __forceinline__ double octave_sum(double x, double y, double z, double acc,
                                             const OctavePlan &P, const NoiseTables &T,
                                             bool &bad, F &&on_octave)
{
    const int n_oct = P.n_oct;
    if (n_oct <= 0) return acc;
    const double r0 = P.rough[0];
    const bool dbl = P.doubling && __all(doubles_exactly(x, y, z, r0));
    const double rc = P.rough[P.check_oct];
    bad = bad | !in_domain(skew3(x * rc, y * rc, z * rc));
    acc += P.base;
    float accf = 0.0f;
    auto octave = [&](int o, const Skewed &s) {
        NoiseSel sel;
        const float v = noise3_sum<SEL>(s.x, s.y, s.z, T, &sel);
        accf = fmaf(v, P.wsumf[o], accf);
        on_octave(o, v, sel);
    };
    if (dbl) {
        Skewed s = skew3(x * r0, y * r0, z * r0);
        for (int o = 0; o < n_oct; ++o) {
            octave(o, s);
            s.x += s.x;
            s.y += s.y;
            s.z += s.z;
        }
    } else {
        for (int o = 0; o < n_oct; ++o) {
            const double r = P.rough[o];
            octave(o, skew3(x * r, y * r, z * r));
        }
    }
    return acc + (double)accf;
}
#endif

}  // namespace nx
