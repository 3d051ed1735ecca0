/*
 * nx_oracle.c -- FP64 CPU restatement of the reference terrain hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  The product
 * library under nixis_amd/ never includes, links or calls this file.
 *
 * Build: see oracle/Makefile (-O2 -fopenmp -ffp-contract=off, no fast-math).
 * Each function cites the reference file:line it restates.  Operation order
 * follows the Python source left to right, which is the order numba lowers
 * to IEEE double operations.
 *
 * Pinning (what checks each function beyond "it reads like the source"):
 *   ora_init            -- permutation / pgi structure invariants; the LCG
 *                          reading (numba int32(int32) truncation in over())
 *                          is parity unpinned: no numba run is possible here.
 *   ora_noise3d         -- a structurally different numpy derivation
 *                          (region logic reduced to lattice-offset vectors and
 *                          one generic delta formula, tests/noise_offsets.py)
 *                          on 10,000 random and crafted points: selection
 *                          identical, values within 1e-13; the 512 period of
 *                          the skewed lattice (hash wraps at 256); |n| <= 1
 *                          and non-degenerate statistics; every region and
 *                          extra-vertex layout exercised
 *                          (tests/test_oracle_noise.py).  No lattice-sum
 *                          identity holds: about 4% of points have a lattice
 *                          point with positive attenuation outside the set
 *                          OpenSimplex selects, so none is claimed.
 *   ora_erosion_*       -- hand-worked values on a four-vertex graph (higher,
 *                          lower and level neighbours, the capacity clamp,
 *                          negative water), Jacobi reads of old values only,
 *                          per-vertex h+s conservation and the flat-field
 *                          closed form for water on small icospheres
 *                          (tests/test_oracle_erosion.py).
 *   ora_rescale         -- closed form on affine inputs; the min/max of an
 *                          array with a NaN follow numba's np.min/np.max
 *                          source (the first NaN, for both), checked by
 *                          its consequences for every form
 *                          (tests/test_oracle_post.py).
 */
#define _GNU_SOURCE /* sincos */
#include "nx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* opensimplex.py:26-27,32 */
static const double STRETCH_3D = -1.0 / 6;
static const double SQUISH_3D = 1.0 / 3;
static const double NORM_3D = 103;

/* opensimplex.py:51-60 -- the 24 gradient vectors of a rhombicuboctahedron
 * (magnitude 11 on one axis, 4 on the other two). */
static const int GRAD3[72] = {
    -11, 4,  4,  -4, 11,  4,  -4, 4,  11,  11, 4,  4,  4, 11,  4,  4, 4,  11,
    -11, -4, 4,  -4, -11, 4,  -4, -4, 11,  11, -4, 4,  4, -11, 4,  4, -4, 11,
    -11, 4,  -4, -4, 11,  -4, -4, 4,  -11, 11, 4,  -4, 4, 11,  -4, 4, 4,  -11,
    -11, -4, -4, -4, -11, -4, -4, -4, -11, 11, -4, -4, 4, -11, -4, 4, -4, -11,
};

int ora_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------- init -- */

/* opensimplex.py:90-107.  over() is compiled with the explicit signature
 * int32(int32) (opensimplex.py:110-112); numba converts the int64 argument
 * to int32 (truncation to the low 32 bits, signed), so every LCG step is
 * s <- int32(s * 6364136223846793005 + 1442695040888963407) with the int64
 * products wrapping mod 2^64.  r = (s + 31) mod (i + 1) uses Python's floored
 * modulo (numba implements it for signed integers), so r >= 0. */
static int64_t lcg_step(int64_t s)
{
    uint64_t t = (uint64_t)s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (int64_t)(int32_t)(uint32_t)t;
}

void ora_init(int64_t seed, int32_t *perm, int32_t *pgi)
{
    int64_t source[256];
    for (int i = 0; i < 256; ++i) source[i] = i;
    int64_t s = seed;
    s = lcg_step(s);
    s = lcg_step(s);
    s = lcg_step(s);
    for (int i = 255; i >= 0; --i) {
        s = lcg_step(s);
        int64_t m = i + 1;
        int64_t r = (s + 31) % m;
        if (r < 0) r += m; /* floored modulo */
        perm[i] = (int32_t)source[r];
        /* int((perm[i] % (72 / 3)) * 3): float modulo of a non-negative int */
        pgi[i] = (int32_t)(fmod((double)perm[i], 24.0) * 3.0);
        source[r] = source[i];
    }
}

/* ------------------------------------------------------------- noise3d -- */

/* opensimplex.py:18-21 */
static inline int64_t fastfloor(double x)
{
    int64_t xi = (int64_t)x;
    return (x < (double)xi) ? xi - 1 : xi;
}

/* opensimplex.py:123-130 */
static inline int grad_index(int64_t xsb, int64_t ysb, int64_t zsb,
                             const int32_t *perm, const int32_t *pgi)
{
    return pgi[(perm[(perm[xsb & 0xFF] + ysb) & 0xFF] + zsb) & 0xFF];
}

static inline double extrapolate3d(int64_t xsb, int64_t ysb, int64_t zsb,
                                   double dx, double dy, double dz,
                                   const int32_t *perm, const int32_t *pgi)
{
    int idx = grad_index(xsb, ysb, zsb, perm, pgi);
    return (double)GRAD3[idx] * dx + (double)GRAD3[idx + 1] * dy +
           (double)GRAD3[idx + 2] * dz;
}

/* One lattice contribution (the repeated block of opensimplex.py, e.g.
 * 389-394): attn = 2 - dx^2 - dy^2 - dz^2; if attn > 0, add attn^4 * ext. */
static inline double contrib(double value, int64_t xs, int64_t ys, int64_t zs,
                             double dx, double dy, double dz,
                             const int32_t *perm, const int32_t *pgi)
{
    double attn = 2 - dx * dx - dy * dy - dz * dz;
    if (attn > 0) {
        attn *= attn;
        value += attn * attn * extrapolate3d(xs, ys, zs, dx, dy, dz, perm, pgi);
    }
    return value;
}

static inline uint32_t sel_mix(uint32_t cs, int64_t xs, int64_t ys,
                               int64_t zs, const int32_t *perm,
                               const int32_t *pgi)
{
    return cs * 31u + (uint32_t)grad_index(xs, ys, zs, perm, pgi);
}

/* opensimplex.py:266-759, statement by statement. */
double ora_noise3d(double x, double y, double z, const int32_t *perm,
                   const int32_t *pgi, int32_t *sel)
{
    /* 271-275 */
    double stretch_offset = (x + y + z) * STRETCH_3D;
    double xs = x + stretch_offset;
    double ys = y + stretch_offset;
    double zs = z + stretch_offset;
    /* 277-280 */
    int64_t xsb = fastfloor(xs);
    int64_t ysb = fastfloor(ys);
    int64_t zsb = fastfloor(zs);
    /* 282-286: int64 sum, then float64 multiply */
    double squish_offset = (double)(xsb + ysb + zsb) * SQUISH_3D;
    double xb = (double)xsb + squish_offset;
    double yb = (double)ysb + squish_offset;
    double zb = (double)zsb + squish_offset;
    /* 288-294 */
    double xins = xs - (double)xsb;
    double yins = ys - (double)ysb;
    double zins = zs - (double)zsb;
    double in_sum = xins + yins + zins;
    /* 296-299 */
    double dx0 = x - xb;
    double dy0 = y - yb;
    double dz0 = z - zb;

    double dx_ext0, dy_ext0, dz_ext0, dx_ext1, dy_ext1, dz_ext1;
    int64_t xsv_ext0, ysv_ext0, zsv_ext0, xsv_ext1, ysv_ext1, zsv_ext1;
    int region;
    double value = 0;

    if (in_sum <= 1) { /* 302: tetrahedron at (0,0,0) */
        region = 0;
        int a_point = 0x01;
        double a_score = xins;
        int b_point = 0x02;
        double b_score = yins;
        if (a_score >= b_score && zins > b_score) {
            b_score = zins;
            b_point = 0x04;
        } else if (a_score < b_score && zins > a_score) {
            a_score = zins;
            a_point = 0x04;
        }
        double wins = 1 - in_sum;
        if (wins > a_score || wins > b_score) { /* 317 */
            int c = (b_score > a_score) ? b_point : a_point;
            if ((c & 0x01) == 0) {
                xsv_ext0 = xsb - 1;
                xsv_ext1 = xsb;
                dx_ext0 = dx0 + 1;
                dx_ext1 = dx0;
            } else {
                xsv_ext0 = xsv_ext1 = xsb + 1;
                dx_ext0 = dx_ext1 = dx0 - 1;
            }
            if ((c & 0x02) == 0) {
                ysv_ext0 = ysv_ext1 = ysb;
                dy_ext0 = dy_ext1 = dy0;
                if ((c & 0x01) == 0) {
                    ysv_ext1 -= 1;
                    dy_ext1 += 1;
                } else {
                    ysv_ext0 -= 1;
                    dy_ext0 += 1;
                }
            } else {
                ysv_ext0 = ysv_ext1 = ysb + 1;
                dy_ext0 = dy_ext1 = dy0 - 1;
            }
            if ((c & 0x04) == 0) {
                zsv_ext0 = zsb;
                zsv_ext1 = zsb - 1;
                dz_ext0 = dz0;
                dz_ext1 = dz0 + 1;
            } else {
                zsv_ext0 = zsv_ext1 = zsb + 1;
                dz_ext0 = dz_ext1 = dz0 - 1;
            }
        } else { /* 353 */
            int c = (a_point | b_point);
            if ((c & 0x01) == 0) {
                xsv_ext0 = xsb;
                xsv_ext1 = xsb - 1;
                dx_ext0 = dx0 - 2 * SQUISH_3D;
                dx_ext1 = dx0 + 1 - SQUISH_3D;
            } else {
                xsv_ext0 = xsv_ext1 = xsb + 1;
                dx_ext0 = dx0 - 1 - 2 * SQUISH_3D;
                dx_ext1 = dx0 - 1 - SQUISH_3D;
            }
            if ((c & 0x02) == 0) {
                ysv_ext0 = ysb;
                ysv_ext1 = ysb - 1;
                dy_ext0 = dy0 - 2 * SQUISH_3D;
                dy_ext1 = dy0 + 1 - SQUISH_3D;
            } else {
                ysv_ext0 = ysv_ext1 = ysb + 1;
                dy_ext0 = dy0 - 1 - 2 * SQUISH_3D;
                dy_ext1 = dy0 - 1 - SQUISH_3D;
            }
            if ((c & 0x04) == 0) {
                zsv_ext0 = zsb;
                zsv_ext1 = zsb - 1;
                dz_ext0 = dz0 - 2 * SQUISH_3D;
                dz_ext1 = dz0 + 1 - SQUISH_3D;
            } else {
                zsv_ext0 = zsv_ext1 = zsb + 1;
                dz_ext0 = dz0 - 1 - 2 * SQUISH_3D;
                dz_ext1 = dz0 - 1 - SQUISH_3D;
            }
        }
        /* 386-416 */
        value = contrib(value, xsb + 0, ysb + 0, zsb + 0, dx0, dy0, dz0, perm, pgi);
        double dx1 = dx0 - 1 - SQUISH_3D;
        double dy1 = dy0 - 0 - SQUISH_3D;
        double dz1 = dz0 - 0 - SQUISH_3D;
        value = contrib(value, xsb + 1, ysb + 0, zsb + 0, dx1, dy1, dz1, perm, pgi);
        double dx2 = dx0 - 0 - SQUISH_3D;
        double dy2 = dy0 - 1 - SQUISH_3D;
        double dz2 = dz1;
        value = contrib(value, xsb + 0, ysb + 1, zsb + 0, dx2, dy2, dz2, perm, pgi);
        double dx3 = dx2;
        double dy3 = dy1;
        double dz3 = dz0 - 1 - SQUISH_3D;
        value = contrib(value, xsb + 0, ysb + 0, zsb + 1, dx3, dy3, dz3, perm, pgi);
        if (sel) {
            uint32_t cs = 0;
            cs = sel_mix(cs, xsb, ysb, zsb, perm, pgi);
            cs = sel_mix(cs, xsb + 1, ysb, zsb, perm, pgi);
            cs = sel_mix(cs, xsb, ysb + 1, zsb, perm, pgi);
            cs = sel_mix(cs, xsb, ysb, zsb + 1, perm, pgi);
            sel[4] = (int32_t)cs;
        }
    } else if (in_sum >= 2) { /* 417: tetrahedron at (1,1,1) */
        region = 1;
        int a_point = 0x06;
        double a_score = xins;
        int b_point = 0x05;
        double b_score = yins;
        if (a_score <= b_score && zins < b_score) {
            b_score = zins;
            b_point = 0x03;
        } else if (a_score > b_score && zins < a_score) {
            a_score = zins;
            a_point = 0x03;
        }
        double wins = 3 - in_sum;
        if (wins < a_score || wins < b_score) { /* 434 */
            int c = (b_score < a_score) ? b_point : a_point;
            if ((c & 0x01) != 0) {
                xsv_ext0 = xsb + 2;
                xsv_ext1 = xsb + 1;
                dx_ext0 = dx0 - 2 - 3 * SQUISH_3D;
                dx_ext1 = dx0 - 1 - 3 * SQUISH_3D;
            } else {
                xsv_ext0 = xsv_ext1 = xsb;
                dx_ext0 = dx_ext1 = dx0 - 3 * SQUISH_3D;
            }
            if ((c & 0x02) != 0) {
                ysv_ext0 = ysv_ext1 = ysb + 1;
                dy_ext0 = dy_ext1 = dy0 - 1 - 3 * SQUISH_3D;
                if ((c & 0x01) != 0) {
                    ysv_ext1 += 1;
                    dy_ext1 -= 1;
                } else {
                    ysv_ext0 += 1;
                    dy_ext0 -= 1;
                }
            } else {
                ysv_ext0 = ysv_ext1 = ysb;
                dy_ext0 = dy_ext1 = dy0 - 3 * SQUISH_3D;
            }
            if ((c & 0x04) != 0) {
                zsv_ext0 = zsb + 1;
                zsv_ext1 = zsb + 2;
                dz_ext0 = dz0 - 1 - 3 * SQUISH_3D;
                dz_ext1 = dz0 - 2 - 3 * SQUISH_3D;
            } else {
                zsv_ext0 = zsv_ext1 = zsb;
                dz_ext0 = dz_ext1 = dz0 - 3 * SQUISH_3D;
            }
        } else { /* 470 */
            int c = (a_point & b_point);
            if ((c & 0x01) != 0) {
                xsv_ext0 = xsb + 1;
                xsv_ext1 = xsb + 2;
                dx_ext0 = dx0 - 1 - SQUISH_3D;
                dx_ext1 = dx0 - 2 - 2 * SQUISH_3D;
            } else {
                xsv_ext0 = xsv_ext1 = xsb;
                dx_ext0 = dx0 - SQUISH_3D;
                dx_ext1 = dx0 - 2 * SQUISH_3D;
            }
            if ((c & 0x02) != 0) {
                ysv_ext0 = ysb + 1;
                ysv_ext1 = ysb + 2;
                dy_ext0 = dy0 - 1 - SQUISH_3D;
                dy_ext1 = dy0 - 2 - 2 * SQUISH_3D;
            } else {
                ysv_ext0 = ysv_ext1 = ysb;
                dy_ext0 = dy0 - SQUISH_3D;
                dy_ext1 = dy0 - 2 * SQUISH_3D;
            }
            if ((c & 0x04) != 0) {
                zsv_ext0 = zsb + 1;
                zsv_ext1 = zsb + 2;
                dz_ext0 = dz0 - 1 - SQUISH_3D;
                dz_ext1 = dz0 - 2 - 2 * SQUISH_3D;
            } else {
                zsv_ext0 = zsv_ext1 = zsb;
                dz_ext0 = dz0 - SQUISH_3D;
                dz_ext1 = dz0 - 2 * SQUISH_3D;
            }
        }
        /* 503-534 */
        double dx3 = dx0 - 1 - 2 * SQUISH_3D;
        double dy3 = dy0 - 1 - 2 * SQUISH_3D;
        double dz3 = dz0 - 0 - 2 * SQUISH_3D;
        value = contrib(value, xsb + 1, ysb + 1, zsb + 0, dx3, dy3, dz3, perm, pgi);
        double dx2 = dx3;
        double dy2 = dy0 - 0 - 2 * SQUISH_3D;
        double dz2 = dz0 - 1 - 2 * SQUISH_3D;
        value = contrib(value, xsb + 1, ysb + 0, zsb + 1, dx2, dy2, dz2, perm, pgi);
        double dx1 = dx0 - 0 - 2 * SQUISH_3D;
        double dy1 = dy3;
        double dz1 = dz2;
        value = contrib(value, xsb + 0, ysb + 1, zsb + 1, dx1, dy1, dz1, perm, pgi);
        double ex0 = dx0 - 1 - 3 * SQUISH_3D;
        double ey0 = dy0 - 1 - 3 * SQUISH_3D;
        double ez0 = dz0 - 1 - 3 * SQUISH_3D;
        value = contrib(value, xsb + 1, ysb + 1, zsb + 1, ex0, ey0, ez0, perm, pgi);
        if (sel) {
            uint32_t cs = 0;
            cs = sel_mix(cs, xsb + 1, ysb + 1, zsb, perm, pgi);
            cs = sel_mix(cs, xsb + 1, ysb, zsb + 1, perm, pgi);
            cs = sel_mix(cs, xsb, ysb + 1, zsb + 1, perm, pgi);
            cs = sel_mix(cs, xsb + 1, ysb + 1, zsb + 1, perm, pgi);
            sel[4] = (int32_t)cs;
        }
    } else { /* 535: octahedron in between */
        region = 2;
        double a_score, b_score, score;
        int a_point, b_point, a_far, b_far;
        double p1 = xins + yins;
        if (p1 > 1) {
            a_score = p1 - 1;
            a_point = 0x03;
            a_far = 1;
        } else {
            a_score = 1 - p1;
            a_point = 0x04;
            a_far = 0;
        }
        double p2 = xins + zins;
        if (p2 > 1) {
            b_score = p2 - 1;
            b_point = 0x05;
            b_far = 1;
        } else {
            b_score = 1 - p2;
            b_point = 0x02;
            b_far = 0;
        }
        double p3 = yins + zins;
        if (p3 > 1) {
            score = p3 - 1;
            if (a_score <= b_score && a_score < score) {
                a_point = 0x06;
                a_far = 1;
            } else if (a_score > b_score && b_score < score) {
                b_point = 0x06;
                b_far = 1;
            }
        } else {
            score = 1 - p3;
            if (a_score <= b_score && a_score < score) {
                a_point = 0x01;
                a_far = 0;
            } else if (a_score > b_score && b_score < score) {
                b_point = 0x01;
                b_far = 0;
            }
        }
        if (a_far == b_far) { /* 577 */
            if (a_far) {
                dx_ext0 = dx0 - 1 - 3 * SQUISH_3D;
                dy_ext0 = dy0 - 1 - 3 * SQUISH_3D;
                dz_ext0 = dz0 - 1 - 3 * SQUISH_3D;
                xsv_ext0 = xsb + 1;
                ysv_ext0 = ysb + 1;
                zsv_ext0 = zsb + 1;
                int c = (a_point & b_point);
                if ((c & 0x01) != 0) {
                    dx_ext1 = dx0 - 2 - 2 * SQUISH_3D;
                    dy_ext1 = dy0 - 2 * SQUISH_3D;
                    dz_ext1 = dz0 - 2 * SQUISH_3D;
                    xsv_ext1 = xsb + 2;
                    ysv_ext1 = ysb;
                    zsv_ext1 = zsb;
                } else if ((c & 0x02) != 0) {
                    dx_ext1 = dx0 - 2 * SQUISH_3D;
                    dy_ext1 = dy0 - 2 - 2 * SQUISH_3D;
                    dz_ext1 = dz0 - 2 * SQUISH_3D;
                    xsv_ext1 = xsb;
                    ysv_ext1 = ysb + 2;
                    zsv_ext1 = zsb;
                } else {
                    dx_ext1 = dx0 - 2 * SQUISH_3D;
                    dy_ext1 = dy0 - 2 * SQUISH_3D;
                    dz_ext1 = dz0 - 2 - 2 * SQUISH_3D;
                    xsv_ext1 = xsb;
                    ysv_ext1 = ysb;
                    zsv_ext1 = zsb + 2;
                }
            } else {
                dx_ext0 = dx0;
                dy_ext0 = dy0;
                dz_ext0 = dz0;
                xsv_ext0 = xsb;
                ysv_ext0 = ysb;
                zsv_ext0 = zsb;
                int c = (a_point | b_point);
                if ((c & 0x01) == 0) {
                    dx_ext1 = dx0 + 1 - SQUISH_3D;
                    dy_ext1 = dy0 - 1 - SQUISH_3D;
                    dz_ext1 = dz0 - 1 - SQUISH_3D;
                    xsv_ext1 = xsb - 1;
                    ysv_ext1 = ysb + 1;
                    zsv_ext1 = zsb + 1;
                } else if ((c & 0x02) == 0) {
                    dx_ext1 = dx0 - 1 - SQUISH_3D;
                    dy_ext1 = dy0 + 1 - SQUISH_3D;
                    dz_ext1 = dz0 - 1 - SQUISH_3D;
                    xsv_ext1 = xsb + 1;
                    ysv_ext1 = ysb - 1;
                    zsv_ext1 = zsb + 1;
                } else {
                    dx_ext1 = dx0 - 1 - SQUISH_3D;
                    dy_ext1 = dy0 - 1 - SQUISH_3D;
                    dz_ext1 = dz0 + 1 - SQUISH_3D;
                    xsv_ext1 = xsb + 1;
                    ysv_ext1 = ysb + 1;
                    zsv_ext1 = zsb - 1;
                }
            }
        } else { /* 642 */
            int c1, c2;
            if (a_far) {
                c1 = a_point;
                c2 = b_point;
            } else {
                c1 = b_point;
                c2 = a_point;
            }
            if ((c1 & 0x01) == 0) {
                dx_ext0 = dx0 + 1 - SQUISH_3D;
                dy_ext0 = dy0 - 1 - SQUISH_3D;
                dz_ext0 = dz0 - 1 - SQUISH_3D;
                xsv_ext0 = xsb - 1;
                ysv_ext0 = ysb + 1;
                zsv_ext0 = zsb + 1;
            } else if ((c1 & 0x02) == 0) {
                dx_ext0 = dx0 - 1 - SQUISH_3D;
                dy_ext0 = dy0 + 1 - SQUISH_3D;
                dz_ext0 = dz0 - 1 - SQUISH_3D;
                xsv_ext0 = xsb + 1;
                ysv_ext0 = ysb - 1;
                zsv_ext0 = zsb + 1;
            } else {
                dx_ext0 = dx0 - 1 - SQUISH_3D;
                dy_ext0 = dy0 - 1 - SQUISH_3D;
                dz_ext0 = dz0 + 1 - SQUISH_3D;
                xsv_ext0 = xsb + 1;
                ysv_ext0 = ysb + 1;
                zsv_ext0 = zsb - 1;
            }
            dx_ext1 = dx0 - 2 * SQUISH_3D;
            dy_ext1 = dy0 - 2 * SQUISH_3D;
            dz_ext1 = dz0 - 2 * SQUISH_3D;
            xsv_ext1 = xsb;
            ysv_ext1 = ysb;
            zsv_ext1 = zsb;
            if ((c2 & 0x01) != 0) {
                dx_ext1 -= 2;
                xsv_ext1 += 2;
            } else if ((c2 & 0x02) != 0) {
                dy_ext1 -= 2;
                ysv_ext1 += 2;
            } else {
                dz_ext1 -= 2;
                zsv_ext1 += 2;
            }
        }
        /* 689-745 */
        double dx1 = dx0 - 1 - SQUISH_3D;
        double dy1 = dy0 - 0 - SQUISH_3D;
        double dz1 = dz0 - 0 - SQUISH_3D;
        value = contrib(value, xsb + 1, ysb + 0, zsb + 0, dx1, dy1, dz1, perm, pgi);
        double dx2 = dx0 - 0 - SQUISH_3D;
        double dy2 = dy0 - 1 - SQUISH_3D;
        double dz2 = dz1;
        value = contrib(value, xsb + 0, ysb + 1, zsb + 0, dx2, dy2, dz2, perm, pgi);
        double dx3 = dx2;
        double dy3 = dy1;
        double dz3 = dz0 - 1 - SQUISH_3D;
        value = contrib(value, xsb + 0, ysb + 0, zsb + 1, dx3, dy3, dz3, perm, pgi);
        double dx4 = dx0 - 1 - 2 * SQUISH_3D;
        double dy4 = dy0 - 1 - 2 * SQUISH_3D;
        double dz4 = dz0 - 0 - 2 * SQUISH_3D;
        value = contrib(value, xsb + 1, ysb + 1, zsb + 0, dx4, dy4, dz4, perm, pgi);
        double dx5 = dx4;
        double dy5 = dy0 - 0 - 2 * SQUISH_3D;
        double dz5 = dz0 - 1 - 2 * SQUISH_3D;
        value = contrib(value, xsb + 1, ysb + 0, zsb + 1, dx5, dy5, dz5, perm, pgi);
        double dx6 = dx0 - 0 - 2 * SQUISH_3D;
        double dy6 = dy4;
        double dz6 = dz5;
        value = contrib(value, xsb + 0, ysb + 1, zsb + 1, dx6, dy6, dz6, perm, pgi);
        if (sel) {
            uint32_t cs = 0;
            cs = sel_mix(cs, xsb + 1, ysb, zsb, perm, pgi);
            cs = sel_mix(cs, xsb, ysb + 1, zsb, perm, pgi);
            cs = sel_mix(cs, xsb, ysb, zsb + 1, perm, pgi);
            cs = sel_mix(cs, xsb + 1, ysb + 1, zsb, perm, pgi);
            cs = sel_mix(cs, xsb + 1, ysb, zsb + 1, perm, pgi);
            cs = sel_mix(cs, xsb, ysb + 1, zsb + 1, perm, pgi);
            sel[4] = (int32_t)cs;
        }
    }
    /* 747-757 */
    value = contrib(value, xsv_ext0, ysv_ext0, zsv_ext0, dx_ext0, dy_ext0, dz_ext0, perm, pgi);
    value = contrib(value, xsv_ext1, ysv_ext1, zsv_ext1, dx_ext1, dy_ext1, dz_ext1, perm, pgi);

    if (sel) {
        uint32_t cs = (uint32_t)sel[4];
        cs = sel_mix(cs, xsv_ext0, ysv_ext0, zsv_ext0, perm, pgi);
        cs = sel_mix(cs, xsv_ext1, ysv_ext1, zsv_ext1, perm, pgi);
        sel[0] = (int32_t)xsb;
        sel[1] = (int32_t)ysb;
        sel[2] = (int32_t)zsb;
        /* region in bits 0-1; ext0 offsets (o+1, 2 bits per axis) in 2-7;
         * ext1 offsets in 8-13. */
        int code = region;
        code |= (int)(xsv_ext0 - xsb + 1) << 2;
        code |= (int)(ysv_ext0 - ysb + 1) << 4;
        code |= (int)(zsv_ext0 - zsb + 1) << 6;
        code |= (int)(xsv_ext1 - xsb + 1) << 8;
        code |= (int)(ysv_ext1 - ysb + 1) << 10;
        code |= (int)(zsv_ext1 - zsb + 1) << 12;
        sel[3] = code;
        sel[4] = (int32_t)cs;
    }
    return value / NORM_3D; /* 759 */
}

/* opensimplex.py:257-263 */
void ora_noisearr3d(const double *x, const double *y, const double *z,
                    int64_t n, const int32_t *perm, const int32_t *pgi,
                    double *out)
{
    for (int64_t i = 0; i < n; ++i)
        out[i] = ora_noise3d(x[i], y[i], z[i], perm, pgi, NULL);
}

/* ------------------------------------------------------------- terrain -- */

/* terrain.py:12-28 */
void ora_sample_noise(const double *verts, int64_t V, const int32_t *perm,
                      const int32_t *pgi, double n_roughness,
                      double n_strength, double radius, double *out,
                      double *raw)
{
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < V; ++v) {
        /* 17: rough_verts = verts * n_roughness (elementwise FP64) */
        double rx = verts[3 * v + 0] * n_roughness;
        double ry = verts[3 * v + 1] * n_roughness;
        double rz = verts[3 * v + 2] * n_roughness;
        double e = ora_noise3d(rx, ry, rz, perm, pgi, NULL);
        if (raw) raw[v] = e;
        /* 28: (elevations + 1) * 0.5 * n_strength * radius */
        out[v] = (e + 1) * 0.5 * n_strength * radius;
    }
}

/* terrain.py:32-59 */
void ora_sample_octaves(const double *verts, int64_t V, double *elevations,
                        const int32_t *perm, const int32_t *pgi,
                        int n_octaves, double n_init_roughness,
                        double n_init_strength, double n_roughness,
                        double n_persistence, double world_radius,
                        double *raw, int32_t *sel)
{
    double n_freq = n_init_roughness;
    double n_amp = n_init_strength;
    for (int o = 0; o < n_octaves; ++o) {
        double rough = n_freq / world_radius;
        double strength = n_amp / world_radius;
#pragma omp parallel for schedule(static)
        for (int64_t v = 0; v < V; ++v) {
            double rx = verts[3 * v + 0] * rough;
            double ry = verts[3 * v + 1] * rough;
            double rz = verts[3 * v + 2] * rough;
            int32_t *sv = sel ? sel + ((int64_t)o * V + v) * 5 : NULL;
            double e = ora_noise3d(rx, ry, rz, perm, pgi, sv);
            if (raw) raw[(int64_t)o * V + v] = e;
            /* 43: elevations += sample_noise(...) */
            elevations[v] += (e + 1) * 0.5 * strength * world_radius;
        }
        n_freq *= n_roughness;  /* 44 */
        n_amp *= n_persistence; /* 45 */
    }
}

/* util.py:110-140, standard mode */
/* np.min(x) and np.max(x) as numba compiles them for a float array (numba
 * 0.54.1 numba/np/arraymath.py:516-531 and the max counterpart at 599-614):
 * the first NaN in array order is returned at once; otherwise the running
 * extreme replaces on a strict comparison.  n >= 1. */
static void np_min_max(const double *x, int64_t n, double *mn, double *mx)
{
    for (int64_t i = 0; i < n; ++i) {
        if (isnan(x[i])) {
            *mn = *mx = x[i];
            return;
        }
    }
    double lo = x[0], hi = x[0];
    for (int64_t i = 1; i < n; ++i) {
        if (x[i] < lo) lo = x[i];
        if (x[i] > hi) hi = x[i];
    }
    *mn = lo;
    *mx = hi;
}

int ora_rescale(const double *x, int64_t n, double lower, double upper,
                double *out)
{
    if (n <= 0) return 0;
    double x_min, x_max;
    np_min_max(x, n, &x_min, &x_max);
    double x_range = x_max - x_min;
    double new_range = upper - lower;
    if (x_range == 0) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < n; ++a)
        out[a] = ((x[a] - x_min) / x_range) * new_range + lower;
    return 0;
}

/* ------------------------------------------------------------- erosion -- */

/* erosion.py:38-40; numba lowers **2 to x*x */
double ora_calc_distance(const double *v0, const double *v1)
{
    double a = v0[0] - v1[0];
    double b = v0[1] - v1[1];
    double c = v0[2] - v1[2];
    return sqrt(a * a + b * b + c * c);
}

/* erosion.py:34-36 */
double ora_calc_slope(double v0, double v1, double dist)
{
    return (v1 - v0) / (dist + 0.00001);
}

/* erosion.py:197-279 */
void ora_erosion_iteration3(const double *verts, const int32_t *nbrs,
                            int64_t V, double *h, double *w, double *s)
{
    const double evaporation = 0.1 / 320;
    const double solubility = 0.01 / 320;
    const double capacity = 0.2 / 320;
    double *hb = (double *)malloc(sizeof(double) * (size_t)V);
    double *wb = (double *)malloc(sizeof(double) * (size_t)V);
    double *sb = (double *)malloc(sizeof(double) * (size_t)V);
    memcpy(hb, h, sizeof(double) * (size_t)V); /* 199-201 */
    memcpy(wb, w, sizeof(double) * (size_t)V);
    memcpy(sb, s, sizeof(double) * (size_t)V);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < V; ++i) {
        double thisvert = h[i];
        double sed_amt = s[i];
        double wat_amt = w[i];
        for (int k = 0; k < 6; ++k) {
            int32_t n = nbrs[i * 6 + k];
            if (n != -1) {
                double compvert = h[n];
                double d = ora_calc_distance(verts + 3 * i, verts + 3 * (int64_t)n);
                double slope = ora_calc_slope(thisvert, compvert, d);
                if (slope > 0) {
                    sed_amt += solubility * w[n];
                    wat_amt += w[n] * d;
                } else if (slope < 0) {
                    sed_amt -= solubility * w[n];
                    wat_amt -= w[n] * d;
                }
            }
        }
        hb[i] -= sed_amt; /* 262-267 */
        sb[i] += sed_amt;
        wb[i] += wat_amt - wat_amt * evaporation;
        if (sb[i] > (capacity * wb[i])) {
            hb[i] += (sb[i] - (capacity * wb[i]));
            sb[i] -= (sb[i] - (capacity * wb[i]));
        }
    }
    memcpy(h, hb, sizeof(double) * (size_t)V); /* 274-277 */
    memcpy(w, wb, sizeof(double) * (size_t)V);
    memcpy(s, sb, sizeof(double) * (size_t)V);
    free(hb);
    free(wb);
    free(sb);
}

/* erosion.py:171-192 (snapshot images omitted) */
void ora_erode_terrain3(const double *verts, const int32_t *nbrs, int64_t V,
                        double *h, int num_iter, double rain, double *water,
                        double *sediment)
{
    if (num_iter <= 0) num_iter = 1;
    double *w = (double *)calloc((size_t)V, sizeof(double));
    double *s = (double *)calloc((size_t)V, sizeof(double));
    for (int it = 0; it < num_iter; ++it) {
        for (int64_t i = 0; i < V; ++i) w[i] += rain;
        ora_erosion_iteration3(verts, nbrs, V, h, w, s);
    }
    if (water) memcpy(water, w, sizeof(double) * (size_t)V);
    if (sediment) memcpy(sediment, s, sizeof(double) * (size_t)V);
    free(w);
    free(s);
}

/* erosion.py:76-99 */
void ora_erosion_iteration1(const int32_t *nbrs, int64_t V, const double *r,
                            double *w_out)
{
    const double simple_constant = 0.0005;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < V; ++i) {
        double thisvert = r[i];
        double amt = 0;
        for (int k = 0; k < 6; ++k) {
            int32_t n = nbrs[i * 6 + k];
            if (n != -1) {
                if (r[n] > thisvert)
                    amt += simple_constant;
                else if (r[n] < thisvert)
                    amt -= simple_constant;
            }
        }
        w_out[i] = thisvert + amt;
    }
}

/* numba's max(a, 0) follows Python's algorithm: the first argument unless
 * the second is greater, so -0.0 and NaN come back unchanged. */
static inline double dmax0(double a) { return 0.0 > a ? 0.0 : a; }

/* erosion.py:301-381 */
void ora_erosion_iteration4(const double *verts, const int32_t *nbrs,
                            int64_t V, double *h, double *w, double *s)
{
    const double evaporation = 0.1 / 320;
    const double solubility = 0.01 / 320;
    const double capacity = 0.2 / 320;
    double *hb = (double *)malloc(sizeof(double) * (size_t)V);
    double *wb = (double *)malloc(sizeof(double) * (size_t)V);
    double *sb = (double *)malloc(sizeof(double) * (size_t)V);
    memcpy(hb, h, sizeof(double) * (size_t)V);
    memcpy(wb, w, sizeof(double) * (size_t)V);
    memcpy(sb, s, sizeof(double) * (size_t)V);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < V; ++i) {
        double thisvert = h[i];
        double sed_amt = s[i];
        double wat_amt = w[i];
        for (int k = 0; k < 6; ++k) {
            int32_t n = nbrs[i * 6 + k];
            if (n != -1) {
                double compvert = h[n];
                double d = ora_calc_distance(verts + 3 * i, verts + 3 * (int64_t)n);
                double slope = ora_calc_slope(thisvert, compvert, d);
                if (slope > 0) {
                    sed_amt += dmax0(solubility * w[n]);
                    wat_amt += dmax0(w[n] * d);
                } else if (slope < 0) {
                    sed_amt -= dmax0(solubility * w[n]);
                    wat_amt -= dmax0(w[n] * d);
                }
            }
        }
        hb[i] -= dmax0(sed_amt);
        sb[i] += dmax0(sed_amt);
        wb[i] += dmax0(wat_amt - wat_amt * evaporation);
        if (sb[i] > (capacity * wb[i])) {
            hb[i] += dmax0(sb[i] - (capacity * wb[i]));
            sb[i] -= dmax0(sb[i] - (capacity * wb[i]));
        }
    }
    memcpy(h, hb, sizeof(double) * (size_t)V);
    memcpy(w, wb, sizeof(double) * (size_t)V);
    memcpy(s, sb, sizeof(double) * (size_t)V);
    free(hb);
    free(wb);
    free(sb);
}

/* ----------------------------------------------------------- adjacency -- */

/* util.py:580-587 */
static int find_first_i32(int32_t item, const int32_t *vec, int n)
{
    for (int i = 0; i < n; ++i)
        if (vec[i] == item) return i;
    return -1;
}

/* util.py:590-613 */
void ora_build_adjacency(const int32_t *tris, int64_t T, int64_t V,
                         int32_t *adj)
{
    for (int64_t i = 0; i < V * 6; ++i) adj[i] = -1;
    for (int64_t t = 0; t < T; ++t) {
        int32_t a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        int32_t *ra = adj + 6 * (int64_t)a;
        int32_t *rb = adj + 6 * (int64_t)b;
        int32_t *rc = adj + 6 * (int64_t)c;
        /* The reference indexes with find_first's -1 (last slot, Python
         * negative index) when a row is full; a valid icosphere never fills
         * a row beyond its degree. */
        int ia = find_first_i32(-1, ra, 6);
        ra[ia < 0 ? 5 : ia] = b;
        int ib = find_first_i32(-1, rb, 6);
        rb[ib < 0 ? 5 : ib] = c;
        int ic = find_first_i32(-1, rc, 6);
        rc[ic < 0 ? 5 : ic] = a;
    }
}

/* util.py:623-629: first i of arr1 that is in arr2 and is not v */
static int32_t next_vert(int32_t v, const int32_t *arr1, const int32_t *arr2)
{
    for (int a = 0; a < 6; ++a) {
        int32_t i = arr1[a];
        for (int b = 0; b < 6; ++b) {
            if (arr2[b] == i) {
                if (i != v) return i;
                break;
            }
        }
    }
    return -1;
}

/* util.py:638-662.  Membership tests only need each row's set, and each row
 * is read in its original order before it is replaced, so reading the
 * unsorted table gives the result of a race-free run of the reference. */
void ora_sort_adjacency(int32_t *adj, int64_t V)
{
    int32_t *orig = (int32_t *)malloc(sizeof(int32_t) * (size_t)V * 6);
    memcpy(orig, adj, sizeof(int32_t) * (size_t)V * 6);
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < V; ++idx) {
        int deg = idx < 12 ? 5 : 6;
        int32_t visited[6] = {-1, -1, -1, -1, -1, -1};
        int32_t pv = (int32_t)idx;
        int32_t nv = orig[idx * 6 + 0];
        for (int i = 0; i < deg - 1; ++i) {
            visited[i] = nv;
            nv = (nv >= 0) ? next_vert(pv, orig + idx * 6, orig + (int64_t)nv * 6) : -1;
            pv = visited[i];
        }
        visited[deg - 1] = nv;
        for (int k = 0; k < 6; ++k) adj[idx * 6 + k] = visited[k];
    }
    free(orig);
}

/* util.py:110-175 (rescale, every form).  mode: 0 none, 1 'lower',
 * 2 'upper'.  Returns 0; -1 when an executed division has a zero divisor
 * (numba raises ZeroDivisionError); 1 when a mode is given without mid (the
 * reference returns x itself; out is then a copy of x). */
int ora_rescale_ex(const double *x, int64_t n, double lower, double upper, int has_mid,
                   double mid, int mode, int has_umin, double u_min, int has_umax,
                   double u_max, double *out)
{
    if (n <= 0) return 0;
    memcpy(out, x, sizeof(double) * (size_t)n);          /* new_array = np.copy(x) */
    double x_min, x_max;
    np_min_max(x, n, &x_min, &x_max);
    if (has_umin && u_min < x_min) x_min = u_min;       /* 131-134 */
    if (has_umax && u_max > x_max) x_max = u_max;
    int zero = 0;
    if (mode == 0) {
        if (!has_mid) {                                  /* 137-141 */
            const double x_range = x_max - x_min, new_range = upper - lower;
            for (int64_t a = 0; a < n; ++a) {
                if (x_range == 0) zero = 1;
                out[a] = ((x[a] - x_min) / x_range) * new_range + lower;
            }
        } else {                                         /* 143-153 */
            const double x_lower_range = mid - x_min, new_lower_range = mid - lower;
            const double x_upper_range = x_max - mid, new_upper_range = upper - mid;
            for (int64_t a = 0; a < n; ++a) {
                if (x[a] <= mid) {
                    if (x_lower_range == 0) zero = 1;
                    out[a] = ((x[a] - x_min) / x_lower_range) * new_lower_range + lower;
                } else {
                    if (x_upper_range == 0) zero = 1;
                    out[a] = ((x[a] - mid) / x_upper_range) * new_upper_range + mid;
                }
            }
        }
        return zero ? -1 : 0;
    }
    if (!has_mid) return 1;                              /* 156-161 */
    if (mode == 1) {                                     /* 162-167 */
        const double x_range = mid - x_min, new_range = mid - lower;
        for (int64_t a = 0; a < n; ++a)
            if (x[a] <= mid) {
                if (x_range == 0) zero = 1;
                out[a] = ((x[a] - x_min) / x_range) * new_range + lower;
            }
    } else if (mode == 2) {                              /* 169-174 */
        const double x_range = x_max - mid, new_range = upper - mid;
        for (int64_t a = 0; a < n; ++a)
            if (x[a] >= mid) {
                if (x_range == 0) zero = 1;
                out[a] = ((x[a] - mid) / x_range) * new_range + mid;
            }
    }
    return zero ? -1 : 0;
}

/* terrain.py:61-72 */
void ora_make_bool_elevation_mask(const double *h, int64_t n, double level, uint8_t *mask)
{
    for (int64_t i = 0; i < n; ++i) mask[i] = h[i] <= level ? 1 : 0;
}

/* util.py:178-254 (power_rescale), sequentially as the reference's range
 * loops; mode 1 rescales the masked values, mode 0 the unmasked ones.
 * Returns -1 when a rescaled value divides by a zero mask range. */
int ora_power_rescale(const double *x, const uint8_t *mask, int64_t n, int mode, double power,
                      double *out)
{
    if (n <= 0) return 0;
    memcpy(out, x, sizeof(double) * (size_t)n);
    const double temp_upper = 1.0, temp_lower = 0.0;
    double x_min, x_max;
    np_min_max(x, n, &x_min, &x_max);
    double mask_lower = x_max, mask_upper = x_min;
    const double temp_range = temp_upper - temp_lower;
    for (int64_t a = 0; a < n; ++a) {
        const int sel = (mode == 1 && mask[a]) || (mode == 0 && !mask[a]);
        if (sel && x[a] < mask_lower)
            mask_lower = x[a];
        else if (sel && x[a] > mask_upper)
            mask_upper = x[a];
    }
    const double mask_range = mask_upper - mask_lower;
    int zero = 0;
    for (int64_t a = 0; a < n; ++a) {
        const int sel = (mode == 1 && mask[a]) || (mode == 0 && !mask[a]);
        if (!sel) continue;
        if (mask_range == 0) zero = 1;
        double t = ((x[a] - mask_lower) / mask_range) * temp_range + temp_lower;
        t = pow(t, power);
        out[a] = ((t - temp_lower) / temp_range) * mask_range + mask_lower;
    }
    return zero ? -1 : 0;
}

/* util.py:556-566 */
double ora_find_percent_val(double minval, double maxval, double percent)
{
    if (!(0.0 < percent && percent < 100.0)) percent = 50.0;
    return minval + ((maxval - minval) * percent / 100.0);
}

/* nixis.py:330-364: the height chain after noise.  h is replaced by the
 * final heights, mask receives the ocean mask; returns the ocean level. */
double ora_height_chain(double *h, int64_t n, double min_alt, double max_alt,
                        double ocean_percent, uint8_t *mask, int *rc)
{
    double *t = (double *)malloc(sizeof(double) * (size_t)n);
    int r = ora_rescale_ex(h, n, min_alt, max_alt, 0, 0.0, 0, 0, 0.0, 0, 0.0, t);
    double minval = t[0], maxval = t[0];
    for (int64_t i = 1; i < n; ++i) {
        if (t[i] < minval) minval = t[i];
        if (t[i] > maxval) maxval = t[i];
    }
    const double ocean_level = ora_find_percent_val(minval, maxval, ocean_percent);
    ora_make_bool_elevation_mask(t, n, ocean_level, mask);
    int r2 = ora_power_rescale(t, mask, n, 1, 0.5, h);
    int r3 = ora_power_rescale(h, mask, n, 0, 2.0, t);
    for (int64_t i = 0; i < n; ++i) t[i] -= ocean_level;
    int r4 = ora_rescale_ex(t, n, min_alt, max_alt, 1, 0.0, 0, 0, 0.0, 0, 0.0, h);
    free(t);
    *rc = (r || r2 || r3 || r4) ? -1 : 0;
    return ocean_level;
}

/* util.py:291-310 with util.py:80-90.  lat and lon follow the reference's
 * expressions (90 - h * latpercent, -180 + w * lonpercent, each clamped with
 * max); x = (r * cos(lat * (pi/180))) * cos(lon * (pi/180)) and so on.  GCC
 * pairs each sin and cos of one angle into a glibc sincos call, as LLVM does
 * for numba's sin/cos intrinsics; whether numba's build does so is not
 * checked (parity unpinned to the last bit of a few coordinates). */
void ora_make_ll_arr(int64_t width, int64_t height, double radius, double *out)
{
    const double latpercent = 180.0 / (double)(height - 1);
    const double lonpercent = 360.0 / (double)width;
    const double d2r = 3.141592653589793 / 180.0;  /* np.pi / 180 */
#pragma omp parallel for schedule(static)
    for (int64_t h = 0; h < height; ++h) {
        double lat = 90.0 - ((double)h * latpercent);
        if (lat < -90.0) lat = -90.0;
        for (int64_t w = 0; w < width; ++w) {
            double lon = -180.0 + ((double)w * lonpercent);
            if (lon < -180.0) lon = -180.0;
            double *o = out + (h * width + w) * 3;
            o[0] = radius * cos(lat * d2r) * cos(lon * d2r);
            o[1] = radius * cos(lat * d2r) * sin(lon * d2r);
            o[2] = radius * sin(lat * d2r);
        }
    }
}

/* scipy.spatial.KDTree.query, brute force (see the header). */
void ora_knn(const double *pts, int64_t V, const double *q, int64_t nq, int k, double *dists,
             int64_t *nbrs)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t a = 0; a < nq; ++a) {
        double bd[16];
        int64_t bi[16];
        int have = 0;
        const double *p = q + 3 * a;
        for (int64_t v = 0; v < V; ++v) {
            const double *c = pts + 3 * v;
            double s = 0.0, d;
            d = p[0] - c[0];
            s += d * d;
            d = p[1] - c[1];
            s += d * d;
            d = p[2] - c[2];
            s += d * d;
            if (have == k && !(s < bd[k - 1])) continue;
            int j = have < k ? have++ : k - 1;
            while (j > 0 && s < bd[j - 1]) {
                bd[j] = bd[j - 1];
                bi[j] = bi[j - 1];
                --j;
            }
            bd[j] = s;
            bi[j] = v;
        }
        for (int j = 0; j < k; ++j) {
            dists[a * k + j] = j < have ? sqrt(bd[j]) : INFINITY;
            nbrs[a * k + j] = j < have ? bi[j] : V;
        }
    }
}

/* util.py:343-366: sd = sum(d); w_i = 1 / ((d_i + 0.00001) / sd);
 * t = sum(w); iw_i = w_i / t; value = int(sum(colors[n_i] * iw_i)), each sum
 * accumulated from 0 in neighbour order. */
void ora_make_gray_array(int64_t npix, int k, const double *dists, const int64_t *nbrs,
                         const double *colors, int32_t *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npix; ++p) {
        const double *d = dists + p * k;
        const int64_t *n = nbrs + p * k;
        double sd = 0.0;
        for (int i = 0; i < k; ++i) sd += d[i];
        double ws[16], t = 0.0;
        for (int i = 0; i < k; ++i) {
            ws[i] = 1.0 / ((d[i] + 0.00001) / sd);
            t += ws[i];
        }
        double v = 0.0;
        for (int i = 0; i < k; ++i) v += colors[n[i]] * (ws[i] / t);
        out[p] = (int32_t)(int64_t)v;
    }
}

/* ------------------------------------------------------------ insolation -- */

/* Python's max(a, b) / min(a, b): the first argument unless the second is
 * strictly greater / smaller. */
static inline double py_max(double a, double b) { return b > a ? b : a; }
static inline double py_min(double a, double b) { return b < a ? b : a; }

/* util.py:60-76 (xyz2latlon): degrees(asin(min(max(z / r, -1), 1))) and
 * degrees(atan2(y, x)), degrees(v) = v * (180 / pi). */
static inline void ora_xyz2latlon(double x, double y, double z, double r, double *lat,
                                  double *lon)
{
    const double r2d = 180.0 / 3.141592653589793;
    *lat = asin(py_min(py_max(z / r, -1.0), 1.0)) * r2d;
    *lon = atan2(y, x) * r2d;
}

/* climate.py:416-446 for one rotation (one call of sample_insolation). */
static void ora_insolation_step(float *arr, const double *verts, int64_t V, double radius,
                                double rotation, double tilt)
{
    const double pi = 3.141592653589793;
    double sr, cr, st, ct;
    sincos(rotation * pi / 180, &sr, &cr);
    sincos(tilt * pi / 180, &st, &ct);
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < V; ++v) {
        const double x = verts[3 * v], y = verts[3 * v + 1], z = verts[3 * v + 2];
        const double rx = x * cr - y * sr;
        const double ry = x * sr + y * cr;
        const double rz = z;
        const double tx = rx * ct + rz * st;
        const double ty = ry;
        const double tz = rz * ct - rx * st;
        double lat, lon;
        ora_xyz2latlon(tx, ty, tz, radius, &lat, &lon);
        const double val = py_max(cos(fabs(lat) * pi / 180), 0.0) * py_max(cos(lon * pi / 180), 0.0);
        arr[v] = (float)((double)arr[v] + val);
    }
}

void ora_sample_insolation(float *arr, const double *verts, int64_t V, double radius,
                           double rot0, double rot_step, int nsteps, double tilt)
{
    double rotation = rot0;
    for (int i = 0; i < nsteps; ++i) {
        ora_insolation_step(arr, verts, V, radius, rotation, tilt);
        rotation += rot_step;
    }
}

void ora_insolation_slice(double radius, double tilt, float *out)
{
    double verts[181 * 3];
    const double d2r = 3.141592653589793 / 180.0;
    for (int i = -90; i <= 90; ++i) {  /* latlon2xyz(i, 0, radius) at index i mod 181 */
        const int k = i < 0 ? 181 + i : i;
        double s, c;
        sincos((double)i * d2r, &s, &c);
        verts[3 * k] = radius * c * 1.0;   /* cos(0 * (pi/180)) == 1 */
        verts[3 * k + 1] = radius * c * 0.0;
        verts[3 * k + 2] = radius * s;
    }
    for (int k = 0; k < 181; ++k) out[k] = 0.0f;
    ora_sample_insolation(out, verts, 181, radius, -180.0, 360.0 / 360, 360, tilt);
}

void ora_interpolate_insolation(const double *verts, int64_t V, const float *table, float *out,
                                double radius)
{
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < V; ++v) {
        double lat, lon;
        ora_xyz2latlon(verts[3 * v], verts[3 * v + 1], verts[3 * v + 2], radius, &lat, &lon);
        const int64_t lower = (int64_t)floor(lat), upper = (int64_t)ceil(lat);
        const float tl = table[lower < 0 ? 181 + lower : lower];
        const float tu = table[upper < 0 ? 181 + upper : upper];
        if (lower == upper) {
            out[v] = tl;
        } else {
            const float diff = tu - tl;  /* float32 subtraction */
            out[v] = (float)((double)tl + (lat - (double)lower) * ((double)diff / (double)(upper - lower)));
        }
    }
}

/* climate.py:351-412 (assign_surface_temp): h2 = rescale(altitudes, 0,
 * alt_intensity) with the reference's alt_intensity of 0 (climate.py:373-374),
 * then climate.py:408 per vertex; -1 when the altitudes are constant (the
 * rescale's division by zero). */
int ora_assign_surface_temp(const double *verts, const double *altitudes, int64_t V,
                            double radius, double tilt, float *out)
{
    const double pi = 3.141592653589793;
    const double alt_intensity = 0;
    double *h2 = (double *)malloc(sizeof(double) * (size_t)(V > 0 ? V : 1));
    if (!h2) return -2;
    if (ora_rescale(altitudes, V, 0, alt_intensity, h2)) {
        free(h2);
        return -1;
    }
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < V; ++v) {
        double lat, lon;
        ora_xyz2latlon(verts[3 * v], verts[3 * v + 1], verts[3 * v + 2], radius, &lat, &lon);
        const double dist = lat - tilt;
        const double c = py_max(cos(fabs(dist) * pi / 180), 0.0);
        out[v] = (float)(h2[v] > 0 ? c - fabs(h2[v]) + 0.1 : c);
    }
    free(h2);
    return 0;
}

void ora_yearly_insolation(const double *verts, int64_t V, double radius, double axial_tilt,
                           float *out)
{
    const double pi = 3.141592653589793;
    float table[181];
    float *daily = (float *)malloc(sizeof(float) * (size_t)V);
    for (int64_t v = 0; v < V; ++v) out[v] = 0.0f;
    for (int x = 0; x < 360; ++x) {
        const double tilt = sin((double)x * pi / 180) * axial_tilt;  /* climate.py:201 */
        ora_insolation_slice(radius, tilt, table);
        ora_interpolate_insolation(verts, V, table, daily, radius);
        for (int64_t v = 0; v < V; ++v) out[v] = out[v] + daily[v];
    }
    free(daily);
}
