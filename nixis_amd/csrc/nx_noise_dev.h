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
// Layout of the evaluation (branch-free across the three regions):
//   - the six unit-cube corners with one or two bits set (100, 010, 001,
//     110, 101, 011) are evaluated for every point and masked by region
//     (region 0 = tetrahedron at 000 keeps the 1-bit corners, region 1 =
//     tetrahedron at 111 keeps the 2-bit corners, region 2 = octahedron
//     keeps all six);
//   - one "origin" slot holds 000 (region 0) or 111 (region 1);
//   - two slots hold the extra vertices chosen by the region logic, whose
//     lattice offsets and delta constants come from a small table keyed by
//     the region's case and closest-point masks.
// The hash prefixes perm[x], perm[perm[x]+y] are shared between corners.
#pragma once

#include <stdint.h>

#include <cmath>
#include <cstring>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NX_ND __host__ __device__ __forceinline__
#else
#define NX_ND inline  // host build of the same code, for CPU checks
#endif

namespace nx {

// Extra-vertex table.  The region logic of opensimplex.py:302-686 reduces
// to one of seven cases and the bit masks of the closest points; `key`
// (case * 64 + mask) selects the two extra lattice offsets (each component
// in -1..2, packed as offset+1 in two bits: ext0 in bits 0-5, ext1 in bits
// 6-11) and their FP32 delta constants o + (ox + oy + oz) / 3.
//   case 0: tetrahedron at 000, (0,0,0) among the closest two (mask = c)
//   case 1: tetrahedron at 000, otherwise (mask = a | b)
//   case 2: tetrahedron at 111, (1,1,1) among the closest two (mask = c)
//   case 3: tetrahedron at 111, otherwise (mask = a & b)
//   case 4: octahedron, both closest on the (1,1,1) side (mask = a & b)
//   case 5: octahedron, both on the (0,0,0) side (mask = a | b)
//   case 6: octahedron, one on each side (mask = far * 8 + near)
struct ExtEntry {
    int32_t code;
    float c0x, c0y, c0z;
    int32_t pad;
    float c1x, c1y, c1z;
};
constexpr int kExtKeys = 7 * 64;

struct ExtLut {
    ExtEntry e[kExtKeys];
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
    const float S = 1.0f / 3.0f;
    const float t0 = S * (float)(e0[0] + e0[1] + e0[2]);
    const float t1 = S * (float)(e1[0] + e1[1] + e1[2]);
    r.c0x = (float)e0[0] + t0;
    r.c0y = (float)e0[1] + t0;
    r.c0z = (float)e0[2] + t0;
    r.c1x = (float)e1[0] + t1;
    r.c1y = (float)e1[1] + t1;
    r.c1z = (float)e1[2] + t1;
    return r;
}

constexpr ExtLut make_ext_lut()
{
    ExtLut t{};
    for (int k = 0; k < 7; ++k)
        for (int m = 0; m < 64; ++m) t.e[k * 64 + m] = ext_entry(k, m);
    return t;
}

// Tables in LDS.  perm and the gradient index (pgi / 3, 0..23) of every hash
// value are stored as bytes with wrap-around padding: entry i + 1 holds
// table[i mod 256] for i in -1..(kWrap - 2), so every hash index used below
// (lattice coordinate mod 256 plus an offset in -1..2, plus a table value)
// lands in range without masking.  Then the 24 gradient vectors as
// (fp16 gx | fp16 gy << 16, FP32 bits of gz) and the extra-vertex table.
constexpr int kWrap = 520;
struct NoiseTables {
    uint8_t perm[kWrap];
    uint8_t grad[kWrap];
    uint32_t g16[24][2];
    ExtEntry ext[kExtKeys];
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

__device__ __forceinline__ void load_tables(NoiseTables &T, const NoiseTablesPacked &P)
{
    const uint8_t *pb = reinterpret_cast<const uint8_t *>(P.perm);
    const uint8_t *gb = reinterpret_cast<const uint8_t *>(P.grad);
    uint32_t *hw = &T.g16[0][0];
    for (int i = threadIdx.x; i < kWrap; i += blockDim.x) {
        T.perm[i] = pb[(i + 255) & 255];
        T.grad[i] = gb[(i + 255) & 255];
    }
    for (int i = threadIdx.x; i < 48; i += blockDim.x) hw[i] = P.g16[i];
    const float4 *src = reinterpret_cast<const float4 *>(kExtLutDev.e);
    float4 *dst = reinterpret_cast<float4 *>(T.ext);
    for (int i = threadIdx.x; i < kExtKeys * 2; i += blockDim.x) dst[i] = src[i];
}
#endif

// One lattice contribution in FP32: attn = 2 - |d|^2 clamped at 0 (the
// reference adds only when attn > 0) times attn^4 * (g . d), accumulated
// into v.  The fp16 gradient components feed the FMAs directly.
// `cap` is 2 for a selected point (attn <= 2 always, so the clamp is the
// reference's attn > 0 test) and 0 for a point the region excludes.
NX_ND float clamp0(float a, float cap)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(a, 0.0f, cap);
#else
    return std::fmin(std::fmax(a, 0.0f), cap);
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

NX_ND float contrib(float v, float dx, float dy, float dz, int gi, const NoiseTables &T,
                    float cap)
{
    float attn = fmaf(-dz, dz, fmaf(-dy, dy, fmaf(-dx, dx, 2.0f)));
    attn = clamp0(attn, cap);
    const float a2 = attn * attn;
    const uint32_t w0 = T.g16[gi][0], w1 = T.g16[gi][1];
    const float dot = fmaf(half_to_float(w0 & 0xffffu), dx,
                           fmaf(half_to_float(w0 >> 16), dy, float_of_bits(w1) * dz));
    return fmaf(a2 * a2, dot, v);
}

// Hash of lattice point (x, y, z) with x, y, z in -1..257 (a coordinate mod
// 256 plus an offset).
NX_ND int hash_full(const NoiseTables &T, int x, int y, int z)
{
    const int a = T.perm[1 + x];
    const int b = T.perm[1 + a + y];
    return T.grad[1 + b + z];
}

// Selection record for parity tests (same encoding as the oracle):
// sel[0..2] = xsb, ysb, zsb; sel[3] = region | ext0 offsets << 2 | ext1
// offsets << 8 (each offset + 1 in two bits per axis); sel[4] = checksum
// cs = cs * 31 + pgi of every selected lattice point in reference order.
struct NoiseSel {
    int v[5];
};

// Evaluate noise3d at (x, y, z) (already multiplied by the octave's
// frequency).  Returns value / 103 in FP32.  With `check`, `bad` is set when
// the point is outside the exact domain (the value is then meaningless); the
// caller checks the octave with the largest frequency, which bounds the
// others.  Every branch of the reference's region logic is evaluated with
// selects, so a wave never diverges.
template <bool SEL>
NX_ND float noise3(double x, double y, double z, const NoiseTables &T, bool check, bool &bad,
                   NoiseSel *sel)
{
    const double STRETCH = -1.0 / 6;
    const double SQUISH = 1.0 / 3;
    // opensimplex.py:272-299 (FP64, reference operation order)
    const double so = (x + y + z) * STRETCH;
    const double xs = x + so, ys = y + so, zs = z + so;
    if (check)
        bad = bad | !((fabs(xs) < kNoiseCoordLimit) & (fabs(ys) < kNoiseCoordLimit) &
                      (fabs(zs) < kNoiseCoordLimit));
    // floor(xs) is the integer the reference's int(floor(xs)) holds, as a
    // double (exact inside the checked domain |xs| < 2^30), so it serves the
    // FP64 arithmetic directly and the int only the hashing.
    const double fx = floor(xs), fy = floor(ys), fz = floor(zs);
    const int xsb = (int)fx, ysb = (int)fy, zsb = (int)fz;
    const double sq = (fx + fy + fz) * SQUISH;
    const double xins = xs - fx, yins = ys - fy, zins = zs - fz;
    const double in_sum = xins + yins + zins;
    // The deltas to the cell origin, x - (fx + sq) (opensimplex.py:297-299),
    // only feed the FP32 section: xins - (so + sq) is the same value up to
    // FP64 rounding, with one shared addition.
    const double t0 = so + sq;
    const double dx0d = xins - t0, dy0d = yins - t0, dz0d = zins - t0;

    // ---- region logic, FP64 comparisons in the reference's form ----------
    // tetrahedron at (0,0,0): opensimplex.py:302-317
    const bool a1 = (xins >= yins) & (zins > yins);
    const bool a2 = !a1 & (xins < yins) & (zins > xins);
    const double asA = a2 ? zins : xins, bsA = a1 ? zins : yins;
    const int apA = a2 ? 4 : 1, bpA = a1 ? 4 : 2;
    const double wA = 1 - in_sum;
    const bool nearA = (wA > asA) | (wA > bsA);
    const bool bA = bsA > asA;
    const int keyA = nearA ? (bA ? bpA : apA) : 64 + (apA | bpA);
    // tetrahedron at (1,1,1): opensimplex.py:417-434
    const bool b1 = (xins <= yins) & (zins < yins);
    const bool b2 = !b1 & (xins > yins) & (zins < xins);
    const double asB = b2 ? zins : xins, bsB = b1 ? zins : yins;
    const int apB = b2 ? 3 : 6, bpB = b1 ? 3 : 5;
    const double wB = 3 - in_sum;
    const bool nearB = (wB < asB) | (wB < bsB);
    const bool bB = bsB < asB;
    const int keyB = nearB ? 128 + (bB ? bpB : apB) : 192 + (apB & bpB);
    // octahedron: opensimplex.py:537-575
    const double p1 = xins + yins, p2 = xins + zins, p3 = yins + zins;
    const bool f1 = p1 > 1, f2 = p2 > 1, f3 = p3 > 1;
    // p - 1 when p > 1, else 1 - p (opensimplex.py:539-559); 1 - p equals
    // -(p - 1) exactly in IEEE arithmetic, so this is |p - 1|.
    const double asC = fabs(p1 - 1), bsC = fabs(p2 - 1), scC = fabs(p3 - 1);
    const bool repA = (asC <= bsC) & (asC < scC);
    const bool repB = !repA & (asC > bsC) & (bsC < scC);
    const int rp = f3 ? 6 : 1;
    const int apC = repA ? rp : (f1 ? 3 : 4), bpC = repB ? rp : (f2 ? 5 : 2);
    const bool afC = repA ? f3 : f1, bfC = repB ? f3 : f2;
    const int kFF = 256 + (apC & bpC), kNN = 320 + (apC | bpC);
    const int kMix = 384 + (afC ? apC * 8 + bpC : bpC * 8 + apC);
    const int keyC = (afC == bfC) ? (afC ? kFF : kNN) : kMix;
    const int region = in_sum <= 1 ? 0 : (in_sum >= 2 ? 1 : 2);
    const int key = region == 0 ? keyA : (region == 1 ? keyB : keyC);

    // ---- FP32 section -----------------------------------------------------
    const float d0x = (float)dx0d, d0y = (float)dy0d, d0z = (float)dz0d;
    const float S = 1.0f / 3.0f;
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

    const float oneBits = region != 1 ? 2.0f : 0.0f;  // 100, 010, 001
    const float twoBits = region != 0 ? 2.0f : 0.0f;  // 110, 101, 011

    // Corner c: delta = d0 - c - S * popcount(c).
    const float ax = d0x - S, ay = d0y - S, az = d0z - S;
    float v = contrib(0.0f, ax - 1.0f, ay, az, g100, T, oneBits);
    v = contrib(v, ax, ay - 1.0f, az, g010, T, oneBits);
    v = contrib(v, ax, ay, az - 1.0f, g001, T, oneBits);
    const float bx = d0x - 2.0f * S, by = d0y - 2.0f * S, bz = d0z - 2.0f * S;
    v = contrib(v, bx - 1.0f, by - 1.0f, bz, g110, T, twoBits);
    v = contrib(v, bx - 1.0f, by, bz - 1.0f, g101, T, twoBits);
    v = contrib(v, bx, by - 1.0f, bz - 1.0f, g011, T, twoBits);
    // Origin slot: 000 in region 0, 111 in region 1.
    const bool r1 = region == 1;
    const int go = T.grad[1 + (r1 ? p11 + 1 : p00) + zm];
    const float k = r1 ? 2.0f : 0.0f;  // 1 + 3S
    v = contrib(v, d0x - k, d0y - k, d0z - k, go, T, region != 2 ? 2.0f : 0.0f);
    // Extra vertices from the table.
    const ExtEntry &E = T.ext[key];
    const int code = E.code;
    const int ge0 = hash_full(T, xm + (code & 3) - 1, ym + ((code >> 2) & 3) - 1,
                              zm + ((code >> 4) & 3) - 1);
    const int ge1 = hash_full(T, xm + ((code >> 6) & 3) - 1, ym + ((code >> 8) & 3) - 1,
                              zm + ((code >> 10) & 3) - 1);
    v = contrib(v, d0x - E.c0x, d0y - E.c0y, d0z - E.c0z, ge0, T, 2.0f);
    v = contrib(v, d0x - E.c1x, d0y - E.c1y, d0z - E.c1z, ge1, T, 2.0f);

    if (SEL) {
        // Reference evaluation order of the selected points; the checksum
        // accumulates pgi = 3 * gradient index.
        uint32_t cs = 0;
        if (region == 0) {
            const int g000 = T.grad[1 + p00 + zm];
            const int order[4] = {g000, g100, g010, g001};
            for (int i = 0; i < 4; ++i) cs = cs * 31u + 3u * (uint32_t)order[i];
        } else if (region == 1) {
            const int g111 = T.grad[2 + p11 + zm];
            const int order[4] = {g110, g101, g011, g111};
            for (int i = 0; i < 4; ++i) cs = cs * 31u + 3u * (uint32_t)order[i];
        } else {
            const int order[6] = {g100, g010, g001, g110, g101, g011};
            for (int i = 0; i < 6; ++i) cs = cs * 31u + 3u * (uint32_t)order[i];
        }
        cs = cs * 31u + 3u * (uint32_t)ge0;
        cs = cs * 31u + 3u * (uint32_t)ge1;
        sel->v[0] = xsb;
        sel->v[1] = ysb;
        sel->v[2] = zsb;
        sel->v[3] = region | (code & 63) << 2 | ((code >> 6) & 63) << 8;
        sel->v[4] = (int)cs;
    }
    return v * (1.0f / 103.0f);
}

}  // namespace nx
