"""Drop-ins for the reference's climate.py insolation (climate.py:193-201,
416-590) and starting surface temperatures (climate.py:351-412) on the GPU.

Vertex arrays are float64 [V, 3]; results are float32 as in the reference.
The `altitudes` arguments of the daily functions only feed a rescale whose
result the reference does not use; it is kept for its one visible effect, a
ZeroDivisionError on a constant array.
"""
import math

import numpy as np

from . import _lib


def calculate_seasonal_tilt(axial_tilt, degrees):
    """climate.py:193-201: sin(degrees * pi / 180) * axial_tilt."""
    return math.sin(degrees * math.pi / 180) * axial_tilt


def _ws(device):
    import torch
    return torch.empty(_lib.lib().nx_climate_workspace_bytes(), dtype=torch.uint8, device=device)


def _verts_dev(verts):
    import torch
    if isinstance(verts, torch.Tensor):
        return verts.to(torch.float64).reshape(-1, 3).contiguous()
    dev = torch.device("cuda", torch.cuda.current_device())
    return torch.from_numpy(np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3)).to(dev)


def _altitude_check(altitudes):
    if altitudes is not None:
        a = np.asarray(altitudes)
        if a.size and a.min() == a.max():
            raise ZeroDivisionError("rescale: input range is zero")


def sample_insolation_device(arr, verts, radius, rotation, tilt, nsteps=1, rot_step=1.0,
                             workspace=None):
    """climate.py:416-446 for nsteps rotations from `rotation` on device
    tensors: arr float32 [V] accumulates, verts float64 [V, 3]."""
    import torch
    if arr.dtype != torch.float32 or not arr.is_contiguous():
        raise ValueError("arr must be a contiguous float32 tensor")
    v = _verts_dev(verts)
    ws = workspace if workspace is not None else _ws(v.device)
    ctx = _lib.Context.get(v.device)
    _lib.check(_lib.lib().nx_sample_insolation(ctx.handle, _lib.ptr(v), v.shape[0], float(radius),
                                               float(rotation), float(rot_step), int(nsteps),
                                               float(tilt), _lib.ptr(arr), _lib.ptr(ws),
                                               _lib.stream_ptr()))
    return arr


def sample_insolation(arr, verts, radius, rotation, tilt):
    """climate.py:416-446: arr (float32 [V], host) updated in place."""
    import torch
    if arr.dtype != np.float32:
        raise TypeError("arr must be float32")
    v = _verts_dev(verts)
    a = torch.from_numpy(np.ascontiguousarray(arr)).to(v.device)
    sample_insolation_device(a, v, radius, rotation, tilt)
    _lib.Context.get(v.device).sync_check()
    arr[...] = a.cpu().numpy()


def brute_daily_insolation(verts, altitudes, radius, tilt, snapshot=False):
    """climate.py:449-486: 360 one-degree rotations from -180 (the snapshot
    images are not produced)."""
    import torch
    _altitude_check(altitudes)
    v = _verts_dev(verts)
    out = torch.zeros(v.shape[0], dtype=torch.float32, device=v.device)
    sample_insolation_device(out, v, radius, -180.0, tilt, 360, 360.0 / 360)
    _lib.Context.get(v.device).sync_check()
    return out.cpu().numpy()


def calc_instant_insolation(verts, altitudes, radius, rotation, tilt):
    """climate.py:489-498."""
    import torch
    v = _verts_dev(verts)
    out = torch.zeros(v.shape[0], dtype=torch.float32, device=v.device)
    sample_insolation_device(out, v, radius, rotation, tilt)
    _lib.Context.get(v.device).sync_check()
    return out.cpu().numpy()


def calc_insolation_slice(radius, tilt):
    """climate.py:503-533 -> float32 [181] (index = latitude)."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    out = torch.empty(181, dtype=torch.float32, device=dev)
    ctx = _lib.Context.get(dev)
    _lib.check(_lib.lib().nx_insolation_slice(ctx.handle, float(radius), float(tilt),
                                              _lib.ptr(out), _lib.ptr(_ws(dev)),
                                              _lib.stream_ptr()))
    ctx.sync_check()
    return out.cpu().numpy()


def interpolate_insolation(verts, lookup_table, insolation, radius):
    """climate.py:547-573: insolation (float32 [V], host) filled in place."""
    import torch
    v = _verts_dev(verts)
    t = torch.from_numpy(np.ascontiguousarray(lookup_table, dtype=np.float32)).to(v.device)
    out = torch.empty(v.shape[0], dtype=torch.float32, device=v.device)
    ctx = _lib.Context.get(v.device)
    _lib.check(_lib.lib().nx_interpolate_insolation(ctx.handle, _lib.ptr(v), v.shape[0],
                                                    _lib.ptr(t), float(radius), _lib.ptr(out),
                                                    _lib.stream_ptr()))
    ctx.sync_check()
    insolation[...] = out.cpu().numpy()


def daily_insolation_device(verts, radius, tilt, out=None, workspace=None):
    import torch
    v = _verts_dev(verts)
    if out is None:
        out = torch.empty(v.shape[0], dtype=torch.float32, device=v.device)
    ws = workspace if workspace is not None else _ws(v.device)
    ctx = _lib.Context.get(v.device)
    _lib.check(_lib.lib().nx_daily_insolation(ctx.handle, _lib.ptr(v), v.shape[0], float(radius),
                                              float(tilt), _lib.ptr(out), _lib.ptr(ws),
                                              _lib.stream_ptr()))
    return out


def calc_daily_insolation(verts, altitudes, radius, tilt):
    """climate.py:536-544."""
    _altitude_check(altitudes)
    out = daily_insolation_device(verts, radius, tilt)
    _lib.Context.get(out.device).sync_check()
    return out.cpu().numpy()


def assign_surface_temp_device(verts, radius, tilt, out=None):
    """climate.py:351-412 on the device -> float32 [V] tensor (see
    assign_surface_temp)."""
    import torch
    v = _verts_dev(verts)
    if out is None:
        out = torch.empty(v.shape[0], dtype=torch.float32, device=v.device)
    ctx = _lib.Context.get(v.device)
    _lib.check(_lib.lib().nx_assign_surface_temp(ctx.handle, _lib.ptr(v), v.shape[0],
                                                 float(radius), float(tilt), _lib.ptr(out),
                                                 _lib.stream_ptr()))
    return out


def assign_surface_temp(verts, altitudes, radius, tilt):
    """climate.py:351-412: starting surface temperatures, float32 [V].  The
    reference weights altitude by 0 (h2 = rescale(altitudes, 0, 0) is 0), so
    each vertex gets max(cos(|lat - tilt| * pi / 180), 0); the altitudes
    only raise ZeroDivisionError when constant, as its rescale does."""
    _altitude_check(altitudes)
    out = assign_surface_temp_device(verts, radius, tilt)
    _lib.Context.get(out.device).sync_check()
    return out.cpu().numpy()


def yearly_insolation_device(points, radius, axial_tilt, out=None, workspace=None):
    """climate.py:576-590 on the device -> float32 [V] tensor."""
    import torch
    v = _verts_dev(points)
    if out is None:
        out = torch.empty(v.shape[0], dtype=torch.float32, device=v.device)
    ws = workspace if workspace is not None else _ws(v.device)
    ctx = _lib.Context.get(v.device)
    _lib.check(_lib.lib().nx_yearly_insolation(ctx.handle, _lib.ptr(v), v.shape[0], float(radius),
                                               float(axial_tilt), _lib.ptr(out), _lib.ptr(ws),
                                               _lib.stream_ptr()))
    return out


def calc_yearly_insolation(points, height, radius, axial_tilt, snapshot=False):
    """climate.py:576-590 (the snapshot images are not produced)."""
    _altitude_check(height)
    out = yearly_insolation_device(points, radius, axial_tilt)
    _lib.Context.get(out.device).sync_check()
    return out.cpu().numpy()
