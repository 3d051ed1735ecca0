"""Drop-ins for the reference's util.py functions on the terrain path.

* ``rescale`` (util.py:110-140, standard mode) runs on the GPU (device
  min/max reduction + map, bit-identical FP64 arithmetic).
* ``create_mesh`` (util.py:17-50) returns this package's own icosphere
  (meshzoo is not available): same counts V = 10d^2+2, T = 20d^2, the 12
  degree-5 vertices first, unit radius.
* ``build_adjacency`` / ``sort_adjacency`` (util.py:590-613, 638-662) run in
  the C library on the host.
"""
import ctypes

import numpy as np

from . import _lib


def rescale_device(x, lower, upper, out=None, stream=None):
    """Standard-mode rescale of a float64 device tensor; returns (out, minmax)."""
    import torch
    if x.dtype != torch.float64 or not x.is_contiguous():
        raise ValueError("x must be a contiguous float64 tensor")
    if out is None:
        out = torch.empty_like(x)
    minmax = torch.empty(2, dtype=torch.float64, device=x.device)
    ctx = _lib.Context.get(x.device)
    _lib.check(_lib.lib().nx_rescale(ctx.handle, _lib.ptr(x), x.numel(), float(lower),
                                     float(upper), _lib.ptr(out), _lib.ptr(minmax),
                                     _lib.stream_ptr(stream)))
    return out, minmax


def rescale(x, lower, upper, mid=None, mode=None, u_min=None, u_max=None):
    """util.py:110-140 (standard mode: mid, mode, u_min, u_max all None).

    Raises ZeroDivisionError when x is constant, as numba's error model does.
    The split-range and one-sided modes are not on the terrain path.
    """
    if mid is not None or mode is not None or u_min is not None or u_max is not None:
        raise NotImplementedError("only the standard rescale mode is provided")
    import torch
    a = np.ascontiguousarray(x, dtype=np.float64)
    if a.size == 0:
        raise ValueError("zero-size array to reduction operation minimum")
    dev = torch.device("cuda", torch.cuda.current_device())
    t = torch.from_numpy(a.ravel()).to(dev)
    out, _ = rescale_device(t, lower, upper, out=t)
    _lib.Context.get(dev).sync_check()
    return out.cpu().numpy().reshape(a.shape)


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
