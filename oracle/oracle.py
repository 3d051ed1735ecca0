"""ctypes wrapper of the FP64 CPU oracle (oracle/nx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by ``bench.py``'s CPU-baseline leg.  The product package
(``nixis_amd``) never imports it.

The functions mirror the reference call signatures (opensimplex.py, terrain.py,
erosion.py, util.py) on numpy arrays.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libnx_oracle.so")
_lib = None

_d = ctypes.c_double
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_p = ctypes.c_void_p


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ora_init.argtypes = [_i64, _p, _p]
        L.ora_noise3d.argtypes = [_d, _d, _d, _p, _p, _p]
        L.ora_noise3d.restype = _d
        L.ora_noisearr3d.argtypes = [_p, _p, _p, _i64, _p, _p, _p]
        L.ora_sample_noise.argtypes = [_p, _i64, _p, _p, _d, _d, _d, _p, _p]
        L.ora_sample_octaves.argtypes = [_p, _i64, _p, _p, _p, _i32, _d, _d,
                                         _d, _d, _d, _p, _p]
        L.ora_rescale.argtypes = [_p, _i64, _d, _d, _p]
        L.ora_rescale.restype = _i32
        L.ora_calc_distance.argtypes = [_p, _p]
        L.ora_calc_distance.restype = _d
        L.ora_calc_slope.argtypes = [_d, _d, _d]
        L.ora_calc_slope.restype = _d
        L.ora_erosion_iteration3.argtypes = [_p, _p, _i64, _p, _p, _p]
        L.ora_erode_terrain3.argtypes = [_p, _p, _i64, _p, _i32, _d, _p, _p]
        L.ora_erosion_iteration1.argtypes = [_p, _i64, _p, _p]
        L.ora_erosion_iteration4.argtypes = [_p, _p, _i64, _p, _p, _p]
        L.ora_build_adjacency.argtypes = [_p, _i64, _i64, _p]
        L.ora_sort_adjacency.argtypes = [_p, _i64]
        L.ora_num_threads.restype = _i32
        L.ora_rescale_ex.argtypes = [_p, _i64, _d, _d, _i32, _d, _i32, _i32, _d, _i32, _d, _p]
        L.ora_rescale_ex.restype = _i32
        L.ora_make_bool_elevation_mask.argtypes = [_p, _i64, _d, _p]
        L.ora_power_rescale.argtypes = [_p, _p, _i64, _i32, _d, _p]
        L.ora_power_rescale.restype = _i32
        L.ora_find_percent_val.argtypes = [_d, _d, _d]
        L.ora_find_percent_val.restype = _d
        L.ora_height_chain.argtypes = [_p, _i64, _d, _d, _d, _p, ctypes.POINTER(_i32)]
        L.ora_height_chain.restype = _d
        L.ora_make_ll_arr.argtypes = [_i64, _i64, _d, _p]
        L.ora_knn.argtypes = [_p, _i64, _p, _i64, ctypes.c_int, _p, _p]
        L.ora_make_gray_array.argtypes = [_i64, ctypes.c_int, _p, _p, _p, _p]
        L.ora_sample_insolation.argtypes = [_p, _p, _i64, _d, _d, _d, ctypes.c_int, _d]
        L.ora_insolation_slice.argtypes = [_d, _d, _p]
        L.ora_interpolate_insolation.argtypes = [_p, _i64, _p, _p, _d]
        L.ora_yearly_insolation.argtypes = [_p, _i64, _d, _d, _p]
        L.ora_assign_surface_temp.argtypes = [_p, _p, _i64, _d, _d, _p]
        L.ora_assign_surface_temp.restype = _i32
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32a(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def num_threads():
    return lib().ora_num_threads()


# ---- opensimplex.py ------------------------------------------------------

def init(seed=0):
    """opensimplex.py:90-107 -> (perm int32[256], pgi int32[256])."""
    perm = np.zeros(256, np.int32)
    pgi = np.zeros(256, np.int32)
    lib().ora_init(int(seed), _ptr(perm), _ptr(pgi))
    return perm, pgi


def noise3d(x, y, z, perm, pgi, with_sel=False):
    """opensimplex.py:266-759 at one point (FP64)."""
    perm = _i32a(perm)
    pgi = _i32a(pgi)
    sel = np.zeros(5, np.int32) if with_sel else None
    v = lib().ora_noise3d(float(x), float(y), float(z), _ptr(perm), _ptr(pgi),
                          _ptr(sel))
    return (v, sel) if with_sel else v


def noisearr3d(x, y, z, perm, pgi):
    """opensimplex.py:257-263."""
    x, y, z = _f64(x), _f64(y), _f64(z)
    out = np.zeros(x.size, np.float64)
    lib().ora_noisearr3d(_ptr(x), _ptr(y), _ptr(z), x.size, _ptr(_i32a(perm)),
                         _ptr(_i32a(pgi)), _ptr(out))
    return out


# ---- terrain.py ----------------------------------------------------------

def sample_noise(verts, perm, pgi, n_roughness=1, n_strength=0.2, radius=1,
                 raw=False):
    """terrain.py:12-28.  With raw=True also returns the noise3d values."""
    verts = _f64(verts)
    V = verts.shape[0]
    out = np.zeros(V, np.float64)
    r = np.zeros(V, np.float64) if raw else None
    lib().ora_sample_noise(_ptr(verts), V, _ptr(_i32a(perm)), _ptr(_i32a(pgi)),
                           float(n_roughness), float(n_strength), float(radius),
                           _ptr(out), _ptr(r))
    return (out, r) if raw else out


def sample_octaves(verts, elevations, perm, pgi, n_octaves=1,
                   n_init_roughness=1.5, n_init_strength=0.4, n_roughness=2.0,
                   n_persistence=0.5, world_radius=1.0, raw=False, sel=False):
    """terrain.py:32-59.  raw -> per-octave noise3d [O,V]; sel -> [O,V,5]."""
    verts = _f64(verts)
    V = verts.shape[0]
    if elevations is None:
        elevations = np.zeros(V, np.float64)
    assert elevations.dtype == np.float64 and elevations.flags.c_contiguous
    r = np.zeros((n_octaves, V), np.float64) if raw else None
    s = np.zeros((n_octaves, V, 5), np.int32) if sel else None
    lib().ora_sample_octaves(_ptr(verts), V, _ptr(elevations), _ptr(_i32a(perm)),
                             _ptr(_i32a(pgi)), int(n_octaves),
                             float(n_init_roughness), float(n_init_strength),
                             float(n_roughness), float(n_persistence),
                             float(world_radius), _ptr(r), _ptr(s))
    extra = []
    if raw:
        extra.append(r)
    if sel:
        extra.append(s)
    return (elevations, *extra) if extra else elevations


# ---- util.py -------------------------------------------------------------

def rescale(x, lower, upper):
    """util.py:110-140, standard mode."""
    x = _f64(x)
    out = np.zeros_like(x)
    rc = lib().ora_rescale(_ptr(x), x.size, float(lower), float(upper), _ptr(out))
    if rc != 0:
        raise ZeroDivisionError("rescale: input range is zero")
    return out


def build_adjacency(tris, V):
    """util.py:590-613 -> int32[V,6], -1 padded."""
    tris = _i32a(tris)
    adj = np.zeros((V, 6), np.int32)
    lib().ora_build_adjacency(_ptr(tris), tris.shape[0], V, _ptr(adj))
    return adj


def sort_adjacency(adj):
    """util.py:638-662, in place."""
    assert adj.dtype == np.int32 and adj.flags.c_contiguous
    lib().ora_sort_adjacency(_ptr(adj), adj.shape[0])


# ---- erosion.py ----------------------------------------------------------

def calc_distance(v0, v1):
    return lib().ora_calc_distance(_ptr(_f64(v0)), _ptr(_f64(v1)))


def calc_slope(v0, v1, dist):
    return lib().ora_calc_slope(float(v0), float(v1), float(dist))


def erosion_iteration3(verts, neighbors, r_buff, wat, sed):
    """erosion.py:197-279; r_buff, wat, sed (float64) updated in place."""
    for a in (r_buff, wat, sed):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    verts, nb = _f64(verts), _i32a(neighbors)
    lib().ora_erosion_iteration3(_ptr(verts), _ptr(nb), nb.shape[0],
                                 _ptr(r_buff), _ptr(wat), _ptr(sed))


def erode_terrain3(nodes, neighbors, heights, num_iter=1, rain=0.3 / 320,
                   return_fields=False):
    """erosion.py:171-192; heights (float64) updated in place."""
    assert heights.dtype == np.float64 and heights.flags.c_contiguous
    verts, nb = _f64(nodes), _i32a(neighbors)
    V = nb.shape[0]
    w = np.zeros(V, np.float64) if return_fields else None
    s = np.zeros(V, np.float64) if return_fields else None
    lib().ora_erode_terrain3(_ptr(verts), _ptr(nb), V, _ptr(heights),
                             int(num_iter), float(rain), _ptr(w), _ptr(s))
    return (w, s) if return_fields else None


def erosion_iteration1(neighbors, r_buff, w_buff):
    """erosion.py:76-99."""
    nb = _i32a(neighbors)
    r = _f64(r_buff)
    assert w_buff.dtype == np.float64 and w_buff.flags.c_contiguous
    lib().ora_erosion_iteration1(_ptr(nb), nb.shape[0], _ptr(r), _ptr(w_buff))
    return w_buff


def erosion_iteration4(verts, neighbors, r_buff, wat, sed):
    """erosion.py:301-381 (without the tie print)."""
    verts, nb = _f64(verts), _i32a(neighbors)
    lib().ora_erosion_iteration4(_ptr(verts), _ptr(nb), nb.shape[0],
                                 _ptr(r_buff), _ptr(wat), _ptr(sed))


# ---- post-processing after noise (util.py, terrain.py, nixis.py) -----------

_MODES = {None: 0, "lower": 1, "upper": 2}


def rescale_ex(x, lower, upper, mid=None, mode=None, u_min=None, u_max=None):
    """util.py:110-175 in every form.  Returns the new array, or x itself
    when a mode is given without mid; ZeroDivisionError on a zero divisor."""
    x = _f64(x)
    out = np.empty_like(x)
    rc = lib().ora_rescale_ex(_ptr(x), x.size, float(lower), float(upper), mid is not None,
                              float(mid or 0.0), _MODES[mode], u_min is not None,
                              float(u_min or 0.0), u_max is not None, float(u_max or 0.0),
                              _ptr(out))
    if rc == -1:
        raise ZeroDivisionError("rescale: zero range")
    if rc == 1:
        return x
    return out


def make_bool_elevation_mask(height, level):
    """terrain.py:61-72"""
    h = _f64(height)
    m = np.zeros(h.size, np.uint8)
    lib().ora_make_bool_elevation_mask(_ptr(h), h.size, float(level), _ptr(m))
    return m.astype(bool)


def power_rescale(x, mask, mode, power):
    """util.py:178-254"""
    x = _f64(x)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.empty_like(x)
    rc = lib().ora_power_rescale(_ptr(x), _ptr(m), x.size, int(mode), float(power), _ptr(out))
    if rc == -1:
        raise ZeroDivisionError("power_rescale: zero mask range")
    return out


def find_percent_val(minval, maxval, percent):
    """util.py:556-566"""
    return lib().ora_find_percent_val(float(minval), float(maxval), float(percent))


def height_chain(height, min_alt=-4000.0, max_alt=8850.0, ocean_percent=55.0):
    """nixis.py:330-364 -> (heights, ocean mask, ocean level)."""
    h = _f64(height).copy()
    m = np.zeros(h.size, np.uint8)
    rc = _i32(0)
    level = lib().ora_height_chain(_ptr(h), h.size, float(min_alt), float(max_alt),
                                   float(ocean_percent), _ptr(m), ctypes.byref(rc))
    if rc.value:
        raise ZeroDivisionError("height chain: zero range")
    return h, m.astype(bool), level


def make_ll_arr(width, height, radius):
    """util.py:291-310 -> float64 [height, width, 3]."""
    out = np.empty((int(height), int(width), 3), np.float64)
    lib().ora_make_ll_arr(int(width), int(height), float(radius), _ptr(out))
    return out


def knn(points, queries, k=3):
    """scipy KDTree.query(queries, k) by brute force -> (dists, nbrs) shaped
    queries.shape[:-1] + (k,)."""
    p = _f64(points).reshape(-1, 3)
    q = _f64(queries)
    shape = q.shape[:-1]
    q = q.reshape(-1, 3)
    d = np.empty((q.shape[0], k), np.float64)
    n = np.empty((q.shape[0], k), np.int64)
    lib().ora_knn(_ptr(p), p.shape[0], _ptr(q), q.shape[0], int(k), _ptr(d), _ptr(n))
    return d.reshape(shape + (k,)), n.reshape(shape + (k,))


def make_gray_array(width, height, dists, nbrs, colors):
    """util.py:343-366 -> int32 [height, width] (before the reference's
    astype to the colors' dtype)."""
    d = _f64(dists)
    k = d.shape[-1]
    n = np.ascontiguousarray(nbrs, dtype=np.int64)
    c = _f64(colors)
    out = np.empty(int(height) * int(width), np.int32)
    lib().ora_make_gray_array(out.size, int(k), _ptr(d), _ptr(n), _ptr(c), _ptr(out))
    return out.reshape(int(height), int(width))


def sample_insolation(arr, verts, radius, rotation, tilt, nsteps=1, rot_step=1.0):
    """climate.py:416-446 (nsteps rotations from `rotation`), float32 arr in place."""
    assert arr.dtype == np.float32 and arr.flags.c_contiguous
    v = _f64(verts)
    lib().ora_sample_insolation(_ptr(arr), _ptr(v), v.shape[0], float(radius), float(rotation),
                                float(rot_step), int(nsteps), float(tilt))
    return arr


def brute_daily_insolation(verts, radius, tilt):
    """climate.py:449-486 without the unused altitude rescale."""
    out = np.zeros(len(verts), np.float32)
    return sample_insolation(out, verts, radius, -180.0, tilt, 360, 360.0 / 360)


def insolation_slice(radius, tilt):
    """climate.py:503-533 -> float32 [181]."""
    out = np.empty(181, np.float32)
    lib().ora_insolation_slice(float(radius), float(tilt), _ptr(out))
    return out


def interpolate_insolation(verts, table, radius):
    """climate.py:547-573 -> float32 [V]."""
    v = _f64(verts)
    t = np.ascontiguousarray(table, dtype=np.float32)
    out = np.empty(v.shape[0], np.float32)
    lib().ora_interpolate_insolation(_ptr(v), v.shape[0], _ptr(t), _ptr(out), float(radius))
    return out


def assign_surface_temp(verts, altitudes, radius, tilt):
    """climate.py:351-412 -> float32 [V]; ZeroDivisionError for constant
    altitudes, as the reference's rescale raises."""
    v = _f64(verts)
    a = _f64(altitudes).reshape(-1)
    out = np.empty(v.shape[0], np.float32)
    rc = lib().ora_assign_surface_temp(_ptr(v), _ptr(a), v.shape[0], float(radius), float(tilt),
                                       _ptr(out))
    if rc == -1:
        raise ZeroDivisionError("rescale: altitude range is zero")
    assert rc == 0
    return out


def yearly_insolation(verts, radius, axial_tilt):
    """climate.py:577-590 -> float32 [V]."""
    v = _f64(verts)
    out = np.empty(v.shape[0], np.float32)
    lib().ora_yearly_insolation(_ptr(v), v.shape[0], float(radius), float(axial_tilt), _ptr(out))
    return out
