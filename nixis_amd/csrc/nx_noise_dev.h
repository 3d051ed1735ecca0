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
//   - two slots hold the extra vertices chosen by the region logic.
// The hash prefixes perm[x], perm[perm[x]+y] are shared between corners.
#pragma once

#include <stdint.h>

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NX_ND __host__ __device__ __forceinline__
#else
#define NX_ND inline  // host build of the same code, for CPU checks
#endif

namespace nx {

// Sign-extended 8-bit field of a packed word.
NX_ND int sbyte(int v, int shift)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sbfe(v, shift, 8);
#else
    return (int)(int8_t)(uint8_t)((unsigned)v >> shift);
#endif
}

// Tables in LDS: perm as bytes, the gradient index (pgi / 3, 0..23) of every
// hash value as bytes, and the 24 gradient vectors packed as signed bytes.
struct NoiseTables {
    uint8_t perm[256];
    uint8_t grad[256];
    int32_t gvec[24];
};

// Host-visible packed form passed in kernel arguments.
struct NoiseTablesPacked {
    uint32_t perm[64];
    uint32_t grad[64];
    int32_t gvec[24];
};

// Exact domain of the FP64 section: |skewed coordinate| < 2^30 keeps every
// lattice integer representable in int32 with margin.
constexpr double kNoiseCoordLimit = 1073741824.0;

#if defined(__HIPCC__)
__device__ __forceinline__ void load_tables(NoiseTables &T, const NoiseTablesPacked &P)
{
    uint32_t *pw = reinterpret_cast<uint32_t *>(T.perm);
    uint32_t *gw = reinterpret_cast<uint32_t *>(T.grad);
    for (int i = threadIdx.x; i < 64; i += blockDim.x) {
        pw[i] = P.perm[i];
        gw[i] = P.grad[i];
    }
    for (int i = threadIdx.x; i < 24; i += blockDim.x) T.gvec[i] = P.gvec[i];
}
#endif

// One lattice contribution in FP32: attn = 2 - |d|^2 clamped at 0 (the
// reference adds only when attn > 0), times attn^4 * (g . d).
NX_ND float contrib(float dx, float dy, float dz, int gi,
                                         const NoiseTables &T, bool on)
{
    float attn = fmaf(-dz, dz, fmaf(-dy, dy, fmaf(-dx, dx, 2.0f)));
    attn = on ? fmaxf(attn, 0.0f) : 0.0f;
    const float a2 = attn * attn;
    const int gv = T.gvec[gi];
    const float gx = (float)sbyte(gv, 0);
    const float gy = (float)sbyte(gv, 8);
    const float gz = (float)sbyte(gv, 16);
    const float dot = fmaf(gz, dz, fmaf(gy, dy, gx * dx));
    return (a2 * a2) * dot;
}

NX_ND int hash_full(const NoiseTables &T, int x, int y, int z)
{
    const int a = T.perm[x & 255];
    const int b = T.perm[(a + y) & 255];
    return T.grad[(b + z) & 255];
}

// Selection record for parity tests (same encoding as the oracle):
// sel[0..2] = xsb, ysb, zsb; sel[3] = region | ext0 offsets << 2 | ext1
// offsets << 8 (each offset + 1 in two bits per axis); sel[4] = checksum
// cs = cs * 31 + pgi of every selected lattice point in reference order.
struct NoiseSel {
    int v[5];
};

// Evaluate noise3d at (x, y, z) (already multiplied by the octave's
// frequency).  Returns value / 103 in FP32.  `bad` is set when the point is
// outside the exact domain (the value is then meaningless).
template <bool SEL>
NX_ND float noise3(double x, double y, double z,
                                        const NoiseTables &T, bool &bad, NoiseSel *sel)
{
    const double STRETCH = -1.0 / 6;
    const double SQUISH = 1.0 / 3;
    // opensimplex.py:272-299 (FP64, reference operation order)
    const double so = (x + y + z) * STRETCH;
    const double xs = x + so, ys = y + so, zs = z + so;
    bad = bad || !(fabs(xs) < kNoiseCoordLimit && fabs(ys) < kNoiseCoordLimit &&
                   fabs(zs) < kNoiseCoordLimit);
    const int xsb = (int)floor(xs), ysb = (int)floor(ys), zsb = (int)floor(zs);
    const double fx = (double)xsb, fy = (double)ysb, fz = (double)zsb;
    const double sq = (fx + fy + fz) * SQUISH;
    const double xins = xs - fx, yins = ys - fy, zins = zs - fz;
    const double in_sum = xins + yins + zins;
    const double dx0d = x - (fx + sq), dy0d = y - (fy + sq), dz0d = z - (fz + sq);

    // Region and extra vertices (opensimplex.py:302-686).  Offsets are
    // relative to (xsb, ysb, zsb), each in -1..2.
    int region;
    int e0x, e0y, e0z, e1x, e1y, e1z;
    if (in_sum <= 1) {
        region = 0;
        int a_point = 1, b_point = 2;
        double a_score = xins, b_score = yins;
        if (a_score >= b_score && zins > b_score) {
            b_score = zins;
            b_point = 4;
        } else if (a_score < b_score && zins > a_score) {
            a_score = zins;
            a_point = 4;
        }
        const double wins = 1 - in_sum;
        if (wins > a_score || wins > b_score) {
            const int c = (b_score > a_score) ? b_point : a_point;
            const int cx = c & 1, cy = c & 2, cz = c & 4;
            e0x = cx ? 1 : -1;
            e1x = cx ? 1 : 0;
            e0y = cy ? 1 : (cx ? -1 : 0);
            e1y = cy ? 1 : (cx ? 0 : -1);
            e0z = cz ? 1 : 0;
            e1z = cz ? 1 : -1;
        } else {
            const int c = a_point | b_point;
            e0x = (c & 1) ? 1 : 0;
            e1x = (c & 1) ? 1 : -1;
            e0y = (c & 2) ? 1 : 0;
            e1y = (c & 2) ? 1 : -1;
            e0z = (c & 4) ? 1 : 0;
            e1z = (c & 4) ? 1 : -1;
        }
    } else if (in_sum >= 2) {
        region = 1;
        int a_point = 6, b_point = 5;
        double a_score = xins, b_score = yins;
        if (a_score <= b_score && zins < b_score) {
            b_score = zins;
            b_point = 3;
        } else if (a_score > b_score && zins < a_score) {
            a_score = zins;
            a_point = 3;
        }
        const double wins = 3 - in_sum;
        if (wins < a_score || wins < b_score) {
            const int c = (b_score < a_score) ? b_point : a_point;
            const int cx = c & 1, cy = c & 2, cz = c & 4;
            e0x = cx ? 2 : 0;
            e1x = cx ? 1 : 0;
            e0y = cy ? (cx ? 1 : 2) : 0;
            e1y = cy ? (cx ? 2 : 1) : 0;
            e0z = cz ? 1 : 0;
            e1z = cz ? 2 : 0;
        } else {
            const int c = a_point & b_point;
            e0x = (c & 1) ? 1 : 0;
            e1x = (c & 1) ? 2 : 0;
            e0y = (c & 2) ? 1 : 0;
            e1y = (c & 2) ? 2 : 0;
            e0z = (c & 4) ? 1 : 0;
            e1z = (c & 4) ? 2 : 0;
        }
    } else {
        region = 2;
        double a_score, b_score;
        int a_point, b_point;
        bool a_far, b_far;
        const double p1 = xins + yins;
        if (p1 > 1) {
            a_score = p1 - 1;
            a_point = 3;
            a_far = true;
        } else {
            a_score = 1 - p1;
            a_point = 4;
            a_far = false;
        }
        const double p2 = xins + zins;
        if (p2 > 1) {
            b_score = p2 - 1;
            b_point = 5;
            b_far = true;
        } else {
            b_score = 1 - p2;
            b_point = 2;
            b_far = false;
        }
        const double p3 = yins + zins;
        if (p3 > 1) {
            const double score = p3 - 1;
            if (a_score <= b_score && a_score < score) {
                a_point = 6;
                a_far = true;
            } else if (a_score > b_score && b_score < score) {
                b_point = 6;
                b_far = true;
            }
        } else {
            const double score = 1 - p3;
            if (a_score <= b_score && a_score < score) {
                a_point = 1;
                a_far = false;
            } else if (a_score > b_score && b_score < score) {
                b_point = 1;
                b_far = false;
            }
        }
        if (a_far == b_far) {
            if (a_far) {
                // (1,1,1) and twice the shared axis
                const int c = a_point & b_point;
                e0x = 1;
                e0y = 1;
                e0z = 1;
                e1x = (c & 1) ? 2 : 0;
                e1y = (!(c & 1) && (c & 2)) ? 2 : 0;
                e1z = (!(c & 1) && !(c & 2)) ? 2 : 0;
            } else {
                // (0,0,0) and a permutation of (-1,1,1) on the omitted axis
                const int c = a_point | b_point;
                e0x = 0;
                e0y = 0;
                e0z = 0;
                const int ax = !(c & 1) ? 0 : (!(c & 2) ? 1 : 2);
                e1x = ax == 0 ? -1 : 1;
                e1y = ax == 1 ? -1 : 1;
                e1z = ax == 2 ? -1 : 1;
            }
        } else {
            const int c1 = a_far ? a_point : b_point;
            const int c2 = a_far ? b_point : a_point;
            const int ax1 = !(c1 & 1) ? 0 : (!(c1 & 2) ? 1 : 2);
            e0x = ax1 == 0 ? -1 : 1;
            e0y = ax1 == 1 ? -1 : 1;
            e0z = ax1 == 2 ? -1 : 1;
            const int ax2 = (c2 & 1) ? 0 : ((c2 & 2) ? 1 : 2);
            e1x = ax2 == 0 ? 2 : 0;
            e1y = ax2 == 1 ? 2 : 0;
            e1z = ax2 == 2 ? 2 : 0;
        }
    }

    // ---- FP32 section ---------------------------------------------------
    const float d0x = (float)dx0d, d0y = (float)dy0d, d0z = (float)dz0d;
    const float S = 1.0f / 3.0f;
    const int xm = xsb & 255, ym = ysb & 255, zm = zsb & 255;

    // Shared hash prefixes for the unit cube.
    const int px0 = T.perm[xm];
    const int px1 = T.perm[(xm + 1) & 255];
    const int p00 = T.perm[(px0 + ym) & 255];
    const int p01 = T.perm[(px0 + ym + 1) & 255];
    const int p10 = T.perm[(px1 + ym) & 255];
    const int p11 = T.perm[(px1 + ym + 1) & 255];
    const int g100 = T.grad[(p10 + zm) & 255];
    const int g010 = T.grad[(p01 + zm) & 255];
    const int g001 = T.grad[(p00 + zm + 1) & 255];
    const int g110 = T.grad[(p11 + zm) & 255];
    const int g101 = T.grad[(p10 + zm + 1) & 255];
    const int g011 = T.grad[(p01 + zm + 1) & 255];

    const bool oneBits = region != 1;  // 100, 010, 001
    const bool twoBits = region != 0;  // 110, 101, 011

    // Corner c: delta = d0 - c - S * popcount(c).
    const float ax = d0x - S, ay = d0y - S, az = d0z - S;  // shared 1-bit part
    float v = contrib(ax - 1.0f, ay, az, g100, T, oneBits);
    v += contrib(ax, ay - 1.0f, az, g010, T, oneBits);
    v += contrib(ax, ay, az - 1.0f, g001, T, oneBits);
    const float bx = d0x - 2.0f * S, by = d0y - 2.0f * S, bz = d0z - 2.0f * S;
    v += contrib(bx - 1.0f, by - 1.0f, bz, g110, T, twoBits);
    v += contrib(bx - 1.0f, by, bz - 1.0f, g101, T, twoBits);
    v += contrib(bx, by - 1.0f, bz - 1.0f, g011, T, twoBits);
    // Origin slot: 000 in region 0, 111 in region 1.
    {
        const bool r1 = region == 1;
        const int go = T.grad[((r1 ? p11 : p00) + zm + (r1 ? 1 : 0)) & 255];
        const float k = r1 ? 2.0f : 0.0f;  // 1 + 3S
        v += contrib(d0x - k, d0y - k, d0z - k, go, T, region != 2);
    }
    // Extra vertices: delta = d0 - o - S * (ox + oy + oz).
    int ge0, ge1;
    {
        const float t0 = S * (float)(e0x + e0y + e0z);
        ge0 = hash_full(T, xm + e0x, ym + e0y, zm + e0z);
        v += contrib(d0x - ((float)e0x + t0), d0y - ((float)e0y + t0),
                     d0z - ((float)e0z + t0), ge0, T, true);
        const float t1 = S * (float)(e1x + e1y + e1z);
        ge1 = hash_full(T, xm + e1x, ym + e1y, zm + e1z);
        v += contrib(d0x - ((float)e1x + t1), d0y - ((float)e1y + t1),
                     d0z - ((float)e1z + t1), ge1, T, true);
    }

    if (SEL) {
        // Reference evaluation order of the selected points; the checksum
        // accumulates pgi = 3 * gradient index.
        uint32_t cs = 0;
        if (region == 0) {
            const int g000 = T.grad[(p00 + zm) & 255];
            const int order[4] = {g000, g100, g010, g001};
            for (int i = 0; i < 4; ++i) cs = cs * 31u + 3u * (uint32_t)order[i];
        } else if (region == 1) {
            const int g111 = T.grad[(p11 + zm + 1) & 255];
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
        sel->v[3] = region | ((e0x + 1) << 2) | ((e0y + 1) << 4) | ((e0z + 1) << 6) |
                    ((e1x + 1) << 8) | ((e1y + 1) << 10) | ((e1z + 1) << 12);
        sel->v[4] = (int)cs;
    }
    return v * (1.0f / 103.0f);
}

}  // namespace nx
