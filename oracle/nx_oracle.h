/*
 * nx_oracle.h -- FP64 CPU restatement of the reference terrain hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing under nixis_amd/ may include, link or
 * call this code: it is the checker that the HIP path is compared against,
 * and (timed on the host cores) the CPU baseline that bench.py reports.
 *
 * Every function restates one reference function in the same IEEE operation
 * order, compiled with -ffp-contract=off and without fast-math, so that each
 * double result equals what numba produces for the reference (numba neither
 * contracts to FMA nor reassociates without fastmath=True).
 */
#ifndef NX_ORACLE_H
#define NX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* opensimplex.py:90-107 (+ over, 110-112).  Parity unpinned for the int32
 * reading of over() (numba's int32(int32) signature): no numba here to
 * confirm it; the tables' structure and the LCG arithmetic are tested. */
void ora_init(int64_t seed, int32_t *perm, int32_t *pgi);

/* opensimplex.py:266-759.  sel (nullable) receives the lattice selection:
 * sel[0..2] = xsb, ysb, zsb; sel[3] = selection code (region + extra vertex
 * offsets); sel[4] = order-sensitive checksum of the gradient indices of the
 * contributing lattice points. */
double ora_noise3d(double x, double y, double z, const int32_t *perm,
                   const int32_t *pgi, int32_t *sel);

/* opensimplex.py:257-263 */
void ora_noisearr3d(const double *x, const double *y, const double *z,
                    int64_t n, const int32_t *perm, const int32_t *pgi,
                    double *out);

/* terrain.py:12-28; verts is [V,3] row-major.  raw (nullable) receives the
 * per-vertex noise3d value before the (e+1)*0.5*s*r weighting. */
void ora_sample_noise(const double *verts, int64_t V, const int32_t *perm,
                      const int32_t *pgi, double n_roughness,
                      double n_strength, double radius, double *out,
                      double *raw);

/* terrain.py:32-59 (without the prints).  elevations is accumulated in place.
 * raw (nullable) is [n_octaves, V]: per-octave noise3d values.  sel (nullable)
 * is [n_octaves, V, 5]: per-octave lattice selection as in ora_noise3d. */
void ora_sample_octaves(const double *verts, int64_t V, double *elevations,
                        const int32_t *perm, const int32_t *pgi,
                        int n_octaves, double n_init_roughness,
                        double n_init_strength, double n_roughness,
                        double n_persistence, double world_radius,
                        double *raw, int32_t *sel);

/* util.py:110-140, standard mode (mid=None, mode=None, no u_min/u_max).
 * Returns 0, or -1 when the input range is zero (numba raises
 * ZeroDivisionError there). */
int ora_rescale(const double *x, int64_t n, double lower, double upper,
                double *out);

/* erosion.py:38-40 and 34-36 */
double ora_calc_distance(const double *v0, const double *v1);
double ora_calc_slope(double v0, double v1, double dist);

/* erosion.py:197-279: one Jacobi step; h, w, s updated in place.
 * nbrs is [V,6] int32 with -1 padding. */
void ora_erosion_iteration3(const double *verts, const int32_t *nbrs,
                            int64_t V, double *h, double *w, double *s);

/* erosion.py:171-192 with the rain amount exposed (reference: 0.3/320).
 * water and sediment (nullable) receive the final fields. */
void ora_erode_terrain3(const double *verts, const int32_t *nbrs, int64_t V,
                        double *h, int num_iter, double rain, double *water,
                        double *sediment);

/* erosion.py:76-99 (v1) and 301-381 (v4, without its tie print).  Not on
 * the product path and not tested: parity unpinned. */
void ora_erosion_iteration1(const int32_t *nbrs, int64_t V, const double *r,
                            double *w_out);
void ora_erosion_iteration4(const double *verts, const int32_t *nbrs,
                            int64_t V, double *h, double *w, double *s);

/* util.py:590-613: adjacency from triangles (serial, as the reference). */
void ora_build_adjacency(const int32_t *tris, int64_t T, int64_t V,
                         int32_t *adj);
/* util.py:638-662: cycle-sort each row; pentagons are rows 0..11 as the
 * reference assumes.  Race-free restatement (reads unsorted rows). */
void ora_sort_adjacency(int32_t *adj, int64_t V);

/* Number of OpenMP threads the oracle will use. */
int ora_num_threads(void);

/* util.py:110-175 (rescale, every form; mode 0 none, 1 'lower', 2 'upper').
 * Returns 0, -1 for a zero divisor, 1 for a mode without mid. */
int ora_rescale_ex(const double *x, int64_t n, double lower, double upper, int has_mid,
                   double mid, int mode, int has_umin, double u_min, int has_umax,
                   double u_max, double *out);
/* terrain.py:61-72 */
void ora_make_bool_elevation_mask(const double *h, int64_t n, double level, uint8_t *mask);
/* util.py:178-254 */
int ora_power_rescale(const double *x, const uint8_t *mask, int64_t n, int mode, double power,
                      double *out);
/* util.py:556-566 */
double ora_find_percent_val(double minval, double maxval, double percent);
/* nixis.py:330-364 */
double ora_height_chain(double *h, int64_t n, double min_alt, double max_alt,
                        double ocean_percent, uint8_t *mask, int *rc);

/* util.py:291-310 (make_ll_arr) with util.py:80-90 (latlon2xyz): out is
 * [height][width][3], row 0 at latitude 90, column 0 at longitude -180. */
void ora_make_ll_arr(int64_t width, int64_t height, double radius, double *out);
/* scipy.spatial.KDTree.query(q, k) as util.py:275-287 and nixis.py:282 use
 * it: the k nearest points, nearest first, distance sqrt of the squared
 * differences summed in x, y, z order.  Brute force; equal distances keep the
 * lower index. */
void ora_knn(const double *pts, int64_t V, const double *q, int64_t nq, int k, double *dists,
             int64_t *nbrs);
/* util.py:343-366 (make_gray_array) over k neighbours per pixel:
 * out[p] = (int)(sum_i colors[nbrs[p][i]] * iw_i). */
void ora_make_gray_array(int64_t npix, int k, const double *dists, const int64_t *nbrs,
                         const double *colors, int32_t *out);

/* climate.py:416-446 (sample_insolation) over nsteps rotations: rotation
 * starts at rot0 and grows by rot_step after each step (climate.py:468-469);
 * each step adds max(cos|lat|, 0) * max(cos lon, 0) of the rotated, tilted
 * vertex to the float32 arr as numba does (double sum, rounded to float). */
void ora_sample_insolation(float *arr, const double *verts, int64_t V, double radius,
                           double rot0, double rot_step, int nsteps, double tilt);
/* climate.py:503-533 (calc_insolation_slice): out[181], index = latitude
 * with negative latitudes at 181 + lat. */
void ora_insolation_slice(double radius, double tilt, float *out);
/* climate.py:351-412 (assign_surface_temp); -1 for constant altitudes */
int ora_assign_surface_temp(const double *verts, const double *altitudes, int64_t V,
                            double radius, double tilt, float *out);
/* climate.py:547-573 (interpolate_insolation) */
void ora_interpolate_insolation(const double *verts, int64_t V, const float *table, float *out,
                                double radius);
/* climate.py:577-590 (calc_yearly_insolation without the snapshot) */
void ora_yearly_insolation(const double *verts, int64_t V, double radius, double axial_tilt,
                           float *out);

#ifdef __cplusplus
}
#endif
#endif
