"""Drop-ins for the reference's util.py functions on the terrain path.

* ``rescale`` (util.py:110-175, every mode) and ``power_rescale``
  (util.py:178-254) run on the GPU (device min/max reductions, a min-scan
  for power_rescale's sequential bounds, FP64 maps in the reference's
  operation order); ``height_chain_device`` runs nixis.py:330-364 after
  noise without leaving the device.
* ``create_mesh`` (util.py:17-50) returns this package's own icosphere
  (meshzoo is not available): same counts V = 10d^2+2, T = 20d^2, the 12
  degree-5 vertices first, unit radius.
* ``build_adjacency`` / ``sort_adjacency`` (util.py:590-613, 638-662) run in
  the C library on the host.
"""
import ctypes

import numpy as np

from . import _lib


_MODES = {None: 0, "lower": 1, "upper": 2}


def _rescale_opts(mid, mode, u_min, u_max):
    if mid is None and mode is None and u_min is None and u_max is None:
        return None
    o = _lib.NxRescaleOpts()
    o.has_mid = mid is not None
    o.mid = float(mid) if mid is not None else 0.0
    o.mode = _MODES[mode]
    o.has_umin = u_min is not None
    o.u_min = float(u_min) if u_min is not None else 0.0
    o.has_umax = u_max is not None
    o.u_max = float(u_max) if u_max is not None else 0.0
    return o


def _f64_device(x, name="x"):
    import torch
    if x.dtype != torch.float64 or not x.is_contiguous() or x.device.type != "cuda":
        raise ValueError(f"{name} must be a contiguous float64 device tensor")


def rescale_device(x, lower, upper, out=None, stream=None, mid=None, mode=None, u_min=None,
                   u_max=None):
    """util.py:110-175 on a float64 device tensor; returns (out, minmax).

    minmax holds the input's (min, max) before u_min / u_max.  out may be x.
    mode is None, 'lower' or 'upper' and needs mid."""
    import torch
    _f64_device(x)
    if mode not in _MODES:
        raise ValueError(f"rescale mode must be None, 'lower' or 'upper', not {mode!r}")
    if mode is not None and mid is None:
        raise ValueError("rescale modes need a mid value")
    if out is None:
        out = torch.empty_like(x)
    minmax = torch.empty(2, dtype=torch.float64, device=x.device)
    ctx = _lib.Context.get(x.device)
    opts = _rescale_opts(mid, mode, u_min, u_max)
    if opts is None:
        rc = _lib.lib().nx_rescale(ctx.handle, _lib.ptr(x), x.numel(), float(lower),
                                   float(upper), _lib.ptr(out), _lib.ptr(minmax),
                                   _lib.stream_ptr(stream))
    else:
        rc = _lib.lib().nx_rescale_ex(ctx.handle, _lib.ptr(x), x.numel(), float(lower),
                                      float(upper), ctypes.byref(opts), _lib.ptr(out),
                                      _lib.ptr(minmax), _lib.stream_ptr(stream))
    _lib.check(rc)
    return out, minmax


def _host_to_device(x):
    import torch
    a = np.ascontiguousarray(x, dtype=np.float64)
    if a.size == 0:
        raise ValueError("zero-size array to reduction operation minimum")
    dev = torch.device("cuda", torch.cuda.current_device())
    return a, dev, torch.from_numpy(a.ravel()).to(dev)


def rescale(x, lower, upper, mid=None, mode=None, u_min=None, u_max=None):
    """util.py:110-175: every form of the reference's rescale, on the GPU.

    As the reference: a mode without mid prints an error and returns x
    itself; a mode other than 'lower' / 'upper' (with mid) returns None; a
    zero range that a value divides by raises ZeroDivisionError (numba's
    error model)."""
    if mode is not None and mid is None:
        print("ERROR: Must supply a middle value to use rescale modes.")
        print("Continuing with unmodified data.")
        return x
    if mode is not None and mode not in ("lower", "upper"):
        return None
    a, dev, t = _host_to_device(x)
    out, _ = rescale_device(t, lower, upper, out=t, mid=mid, mode=mode, u_min=u_min, u_max=u_max)
    _lib.Context.get(dev).sync_check()
    return out.cpu().numpy().reshape(a.shape)


def power_rescale_device(x, mask, mode, power, out=None, stream=None, workspace=None):
    """util.py:178-254 on device tensors: x float64 [n], mask bool/uint8 [n].

    mode 1 rescales the masked values, mode 0 the others.  out may be x."""
    import torch
    _f64_device(x)
    if mask.dtype not in (torch.bool, torch.uint8) or not mask.is_contiguous() \
            or mask.numel() != x.numel() or mask.device != x.device:
        raise ValueError("mask must be a contiguous bool/uint8 tensor of x's length on x's device")
    if out is None:
        out = torch.empty_like(x)
    L = _lib.lib()
    n = x.numel()
    need = L.nx_power_rescale_workspace_bytes(n)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=x.device)
    ctx = _lib.Context.get(x.device)
    _lib.check(L.nx_power_rescale(ctx.handle, _lib.ptr(x), _lib.ptr(mask.view(torch.uint8)), n,
                                  int(mode), float(power), _lib.ptr(out), _lib.ptr(workspace),
                                  _lib.stream_ptr(stream)))
    return out


def power_rescale(x, mask=None, mode=None, power=1.0):
    """util.py:178-254 on the GPU; returns a new array.  Mode None (or any
    value other than 0 and 1) selects nothing and returns a copy, as the
    reference's loops do."""
    import torch
    if mode not in (0, 1):
        return np.array(x, dtype=np.float64, copy=True)
    a, dev, t = _host_to_device(x)
    m = np.ascontiguousarray(mask).astype(np.uint8, copy=False).ravel()
    if m.size != a.size:
        raise ValueError("mask and x differ in length")
    md = torch.from_numpy(m).to(dev)
    out = power_rescale_device(t, md, mode, power, out=t)
    _lib.Context.get(dev).sync_check()
    return out.cpu().numpy().reshape(a.shape)


def find_percent_val(minval, maxval, percent):
    """util.py:556-566: the value `percent` percent of the way from minval to
    maxval; a percent outside (0, 100) is taken as 50 after an error line."""
    if not 0.0 < percent < 100.0:
        print("\n" + "    ERROR: Percent must be between 0 and 100.")
        print("    Defaulting to 50 percent." + "\n")
        percent = 50.0
    return minval + ((maxval - minval) * percent / 100.0)


def find_val_percent(minval, maxval, x):
    """util.py:568-577: where x lies between minval and maxval, in percent
    (an x outside the range prints an error line and is still computed)."""
    if not minval < x < maxval:
        print("\n" + "    ERROR: X must be between minval and maxval.")
    return (x - minval) / (maxval - minval) * 100


def height_chain_device(h, min_alt, max_alt, ocean_percent, stream=None, workspace=None):
    """nixis.py:330-364 in place on the float64 device tensor h (the noise
    heights): rescale to [min_alt, max_alt], ocean level at ocean_percent,
    the ocean mask, power_rescale of the ocean floor (0.5) and of the land
    (2.0), the shift to sea level 0 and the split rescale around 0.

    Returns (h, mask uint8 [n], ocean_level float64 [1]), all on the device;
    nothing is copied to the host."""
    import torch
    _f64_device(h, "h")
    L = _lib.lib()
    n = h.numel()
    need = L.nx_height_chain_workspace_bytes(n)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=h.device)
    mask = torch.empty(n, dtype=torch.uint8, device=h.device)
    level = torch.empty(1, dtype=torch.float64, device=h.device)
    ctx = _lib.Context.get(h.device)
    _lib.check(L.nx_height_chain(ctx.handle, _lib.ptr(h), n, float(min_alt), float(max_alt),
                                 float(ocean_percent), _lib.ptr(mask), _lib.ptr(level),
                                 _lib.ptr(workspace), _lib.stream_ptr(stream)))
    return h, mask, level


def icosphere_counts(divisions):
    V = ctypes.c_int64()
    T = ctypes.c_int64()
    _lib.check(_lib.lib().nx_icosphere_counts(int(divisions), ctypes.byref(V), ctypes.byref(T)))
    return V.value, T.value


def create_mesh(divisions):
    """util.py:17-50 -> (points float64 [V,3], cells int32 [T,3])."""
    V, T = icosphere_counts(divisions)
    pts = np.empty((V, 3), dtype=np.float64)
    cells = np.empty((T, 3), dtype=np.int32)
    _lib.check(_lib.lib().nx_icosphere(int(divisions), _lib.ptr(pts), _lib.ptr(cells)))
    return pts, cells


def create_points(divisions):
    """Vertices only (no triangle list), for noise-only workloads."""
    V, _ = icosphere_counts(divisions)
    pts = np.empty((V, 3), dtype=np.float64)
    _lib.check(_lib.lib().nx_icosphere(int(divisions), _lib.ptr(pts), None))
    return pts


def build_adjacency(triangles):
    """util.py:590-613 -> int32 [V,6] (-1 padded), V = (T + 4) / 2."""
    tris = np.ascontiguousarray(triangles, dtype=np.int32)
    V = (tris.shape[0] + 4) // 2
    adj = np.empty((V, 6), dtype=np.int32)
    _lib.check(_lib.lib().nx_build_adjacency(_lib.ptr(tris), tris.shape[0], V, _lib.ptr(adj)))
    return adj


def sort_adjacency(adj):
    """util.py:638-662: cyclic order of every row, in place."""
    if adj.dtype != np.int32 or not adj.flags.c_contiguous or adj.ndim != 2 or adj.shape[1] != 6:
        raise ValueError("adj must be a C-contiguous int32 [V,6] array")
    _lib.check(_lib.lib().nx_sort_adjacency(_lib.ptr(adj), adj.shape[0]))


def sfc_order(points):
    """Space-filling-curve (3D Morton) order of the vertices: order[new] = old."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    order = np.empty(pts.shape[0], dtype=np.int64)
    _lib.check(_lib.lib().nx_sfc_order(_lib.ptr(pts), pts.shape[0], _lib.ptr(order)))
    return order


def icosphere_with_neighbors(divisions):
    """(points, neighbors) for a d-division icosphere, neighbours cycle-sorted."""
    pts, cells = create_mesh(divisions)
    adj = build_adjacency(cells)
    sort_adjacency(adj)
    return pts, adj


# ---- device-side mesh (nx_*_device): the same arrays, generated on the GPU ----

def create_mesh_device(divisions, device=None, stream=None, cells=True):
    """util.py:17-50 on the device: (points float64 [V,3], cells int32 [T,3]
    or None) tensors equal to create_mesh(divisions) bit for bit."""
    import torch
    V, T = icosphere_counts(divisions)
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    pts = torch.empty((V, 3), dtype=torch.float64, device=device)
    tris = torch.empty((T, 3), dtype=torch.int32, device=device) if cells else None
    ctx = _lib.Context.get(pts.device)
    _lib.check(_lib.lib().nx_icosphere_device(ctx.handle, int(divisions), _lib.ptr(pts),
                                              _lib.ptr(tris), _lib.stream_ptr(stream)))
    return pts, tris


def build_adjacency_device(cells, V=None, stream=None):
    """util.py:590-613 on the device: int32 [V,6] tensor, -1 padded."""
    import torch
    if cells.dtype != torch.int32 or cells.ndim != 2 or cells.shape[1] != 3 or not cells.is_contiguous():
        raise ValueError("cells must be a contiguous int32 [T,3] device tensor")
    T = cells.shape[0]
    V = (T + 4) // 2 if V is None else int(V)
    adj = torch.empty((V, 6), dtype=torch.int32, device=cells.device)
    ws = torch.empty(int(_lib.lib().nx_adjacency_workspace_bytes(V)), dtype=torch.uint8,
                     device=cells.device)
    ctx = _lib.Context.get(cells.device)
    _lib.check(_lib.lib().nx_build_adjacency_device(ctx.handle, _lib.ptr(cells), T, V,
                                                    _lib.ptr(adj), _lib.ptr(ws),
                                                    _lib.stream_ptr(stream)))
    ctx.sync_check(stream)
    return adj


def sort_adjacency_device(adj, stream=None):
    """util.py:638-662 on the device: a new int32 [V,6] tensor in cyclic order."""
    import torch
    if adj.dtype != torch.int32 or adj.ndim != 2 or adj.shape[1] != 6 or not adj.is_contiguous():
        raise ValueError("adj must be a contiguous int32 [V,6] device tensor")
    out = torch.empty_like(adj)
    ctx = _lib.Context.get(adj.device)
    _lib.check(_lib.lib().nx_sort_adjacency_device(ctx.handle, _lib.ptr(adj), _lib.ptr(out),
                                                   adj.shape[0], _lib.stream_ptr(stream)))
    ctx.sync_check(stream)
    return out


def icosphere_with_neighbors_device(divisions, device=None):
    """(points, neighbours) of a d-division icosphere, generated on the device."""
    pts, cells = create_mesh_device(divisions, device)
    return pts, sort_adjacency_device(build_adjacency_device(cells, pts.shape[0]))


def sfc_order_device(points, stream=None):
    """sfc_order on the device: int64 [V] tensor, order[new] = old."""
    import torch
    if points.dtype != torch.float64 or points.ndim != 2 or points.shape[1] != 3 \
            or not points.is_contiguous():
        raise ValueError("points must be a contiguous float64 [V,3] device tensor")
    V = points.shape[0]
    order = torch.empty(V, dtype=torch.int64, device=points.device)
    ws = torch.empty(int(_lib.lib().nx_sfc_workspace_bytes(V)), dtype=torch.uint8,
                     device=points.device)
    ctx = _lib.Context.get(points.device)
    _lib.check(_lib.lib().nx_sfc_order_device(ctx.handle, _lib.ptr(points), V, _lib.ptr(order),
                                              _lib.ptr(ws), _lib.stream_ptr(stream)))
    return order


# ---- equirectangular export (util.py:275-455, nixis.py:270-302) ----------------

class KnnGrid:
    """The k-nearest-neighbour structure that replaces the reference's scipy
    KDTree (util.py:275-287): a uniform grid of cells on the device.
    `query` returns what KDTree.query returns (nearest first, float64
    distances, int64 indices; with an int k of 1 the last axis is dropped);
    equal distances are ordered by index."""

    def __init__(self, points, device=None):
        import torch
        if isinstance(points, torch.Tensor):
            dev = points.device if device is None else torch.device(device)
            p = points.to(dev, torch.float64).reshape(-1, 3).contiguous()
        else:
            dev = torch.device("cuda", torch.cuda.current_device()) if device is None \
                else torch.device(device)
            p = torch.from_numpy(np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)).to(dev)
        self.device = dev
        self.points = p
        self.V = p.shape[0]
        L = _lib.lib()
        self.ws = torch.empty(L.nx_knn_grid_bytes(self.V), dtype=torch.uint8, device=dev)
        self.ctx = _lib.Context.get(dev)
        _lib.check(L.nx_knn_grid_build(self.ctx.handle, _lib.ptr(p), self.V, _lib.ptr(self.ws),
                                       _lib.stream_ptr()))

    def query_device(self, q, k=3):
        """q: float64 device tensor [..., 3] -> (dists [..., k], nbrs [..., k])."""
        import torch
        q = q.to(self.device, torch.float64).contiguous()
        shape = tuple(q.shape[:-1])
        nq = q.numel() // 3
        d = torch.empty(shape + (k,), dtype=torch.float64, device=self.device)
        n = torch.empty(shape + (k,), dtype=torch.int64, device=self.device)
        _lib.check(_lib.lib().nx_knn_query(self.ctx.handle, _lib.ptr(self.ws), self.V,
                                           _lib.ptr(q), nq, int(k), _lib.ptr(d), _lib.ptr(n),
                                           _lib.stream_ptr()))
        return d, n

    def query(self, x, k=3, workers=-1):
        """KDTree.query(x, k, workers) on host arrays; workers is accepted
        and ignored."""
        import torch
        kk = int(k)
        q = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(self.device)
        d, n = self.query_device(q, kk)
        self.ctx.sync_check()
        d, n = d.cpu().numpy(), n.cpu().numpy()
        if isinstance(k, (int, np.integer)) and kk == 1:
            d, n = d[..., 0], n[..., 0]
        return d, n

    def query_pixels_device(self, width, height, radius, k=3):
        """make_ll_arr(width, height, radius) and its query in one pass ->
        (dists [height, width, k], nbrs [height, width, k]) on the device."""
        import torch
        L = _lib.lib()
        ws = torch.empty(L.nx_pixel_workspace_bytes(int(width), int(height)), dtype=torch.uint8,
                         device=self.device)
        d = torch.empty((int(height), int(width), int(k)), dtype=torch.float64, device=self.device)
        n = torch.empty((int(height), int(width), int(k)), dtype=torch.int64, device=self.device)
        _lib.check(L.nx_knn_query_pixels(self.ctx.handle, _lib.ptr(self.ws), self.V, int(width),
                                         int(height), float(radius), int(k), _lib.ptr(d),
                                         _lib.ptr(n), _lib.ptr(ws), _lib.stream_ptr()))
        return d, n


def build_KDTree(points, lf=10):
    """util.py:275-287.  The reference stores the tree in cfg.KDT; this
    returns it (leafsize has no meaning for the grid and is ignored)."""
    return KnnGrid(points)


def make_ll_arr_device(width, height, radius, device=None):
    """util.py:291-310 on the device -> float64 [height, width, 3]."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
    L = _lib.lib()
    ws = torch.empty(L.nx_pixel_workspace_bytes(int(width), int(height)), dtype=torch.uint8,
                     device=dev)
    out = torch.empty((int(height), int(width), 3), dtype=torch.float64, device=dev)
    ctx = _lib.Context.get(dev)
    _lib.check(L.nx_make_ll_arr(ctx.handle, int(width), int(height), float(radius),
                                _lib.ptr(out), _lib.ptr(ws), _lib.stream_ptr()))
    return out


def make_ll_arr(width, height, radius):
    """util.py:291-310 -> float64 [height, width, 3] (host array)."""
    out = make_ll_arr_device(width, height, radius)
    _lib.Context.get(out.device).sync_check()
    return out.cpu().numpy()


def make_gray_array_device(dists, nbrs, colors):
    """util.py:343-366 on device tensors: dists float64 [..., k], nbrs int64
    [..., k], colors float64 [n] -> int32 [...] (before any astype)."""
    import torch
    k = dists.shape[-1]
    if nbrs.shape != dists.shape:
        raise ValueError("dists and nbrs differ in shape")
    d = dists.contiguous()
    n = nbrs.to(torch.int64).contiguous()
    c = colors.to(d.device, torch.float64).contiguous()
    out = torch.empty(tuple(d.shape[:-1]), dtype=torch.int32, device=d.device)
    ctx = _lib.Context.get(d.device)
    _lib.check(_lib.lib().nx_make_gray_array(ctx.handle, out.numel(), int(k), _lib.ptr(d),
                                             _lib.ptr(n), _lib.ptr(c), c.numel(), _lib.ptr(out),
                                             _lib.stream_ptr()))
    return out


def _gray_host(width, height, dists, nbrs, colors):
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    d = torch.from_numpy(np.ascontiguousarray(dists, dtype=np.float64)).to(dev)
    n = torch.from_numpy(np.ascontiguousarray(nbrs, dtype=np.int64)).to(dev)
    c = torch.from_numpy(np.ascontiguousarray(colors, dtype=np.float64).ravel()).to(dev)
    out = make_gray_array_device(d, n, c)
    _lib.Context.get(dev).sync_check()
    return out.cpu().numpy().reshape(int(height), int(width))


def make_gray_array(width, height, dists, nbrs, colors):
    """util.py:343-366: the grayscale map, in the colors' dtype."""
    return _gray_host(width, height, dists, nbrs, colors).astype(np.asarray(colors).dtype)


def make_rgb_array(width, height, dists, nbrs, colors):
    """util.py:313-340: the same value in three channels, in the colors' dtype."""
    g = _gray_host(width, height, dists, nbrs, colors)
    return np.repeat(g[:, :, None], 3, axis=2).astype(np.asarray(colors).dtype)


def build_image_data(colors, width, height, query_data):
    """util.py:369-438.  The reference reads the image size from options.json
    and the query results from cfg.IMG_QUERY_DATA; here they are arguments
    (query_data = (dists, nbrs) as KnnGrid.query returns them)."""
    dists, nbrs = query_data
    result = {}
    if not isinstance(colors, dict):
        print("ERROR: Must pass a dict when saving out texture maps.")
    for key, container in colors.items():
        array, mode = container[0], container[1]
        if array.dtype in ('int8', 'uint8', 'bool_'):
            colors[key] = [rescale(array.astype(np.float64), 0, 255), mode]
        elif array.dtype in ('uint16', 'uint32'):
            pass
        else:
            colors[key] = [rescale(array, 0, 255), mode]
    for key, container in colors.items():
        array, mode = container[0], container[1]
        if mode in ('rgb', 'RGB'):
            pixels = make_rgb_array(width, height, dists, nbrs, array)
        elif mode in ('gray', 'GRAY', 'grey', 'GREY'):
            pixels = make_gray_array(width, height, dists, nbrs, array)
        colors[key] = None
        if array.dtype in ('float16', 'int32', 'uint32', 'float32', 'int64', 'uint64', 'float64'):
            result[key] = pixels.astype('uint8')
        else:
            result[key] = pixels.astype(array.dtype)
    return result


def save_image(data, path, name, fmt="png"):
    """util.py:441-455 with Pillow, as the reference (the format comes from
    options.json there and is an argument here)."""
    import os
    from PIL import Image
    for key, array in data.items():
        Image.fromarray(array).save(os.path.join(path, f"{name}_{key}.{fmt}"))
