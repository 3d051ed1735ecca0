"""Drop-in for the reference's terrain.py noise sampler.

``sample_noise`` (terrain.py:12-28) and ``sample_octaves`` (terrain.py:32-59)
keep their signatures, defaults and in-place semantics; all octaves run in
one fused GPU kernel.  ``*_device`` variants take and return torch tensors
that stay on the GPU.
"""
import numpy as np

from . import _lib


def octave_scalars(n_octaves=1, n_init_roughness=1.5, n_init_strength=0.4,
                   n_roughness=2.0, n_persistence=0.5, world_radius=1.0):
    """Per-octave (rough, strength) as terrain.py:36-45 computes them.

    The frequency and amplitude evolve by repeated multiplication (not pow),
    and each octave calls sample_noise(verts, ..., n_freq / world_radius,
    n_amp / world_radius, world_radius).
    """
    n_octaves = int(n_octaves)
    if n_octaves < 0 or n_octaves > _lib.NX_MAX_OCTAVES:
        raise ValueError(f"n_octaves={n_octaves} outside 0..{_lib.NX_MAX_OCTAVES}")
    n_freq = float(n_init_roughness)
    n_amp = float(n_init_strength)
    R = float(world_radius)
    rough = np.zeros(n_octaves, dtype=np.float64)
    strength = np.zeros(n_octaves, dtype=np.float64)
    for i in range(n_octaves):
        rough[i] = n_freq / R
        strength[i] = n_amp / R
        n_freq *= float(n_roughness)
        n_amp *= float(n_persistence)
    return rough, strength


def sample_octaves_device(verts, elevations, perm, pgi, rough, strength, radius,
                          dbg_noise=None, dbg_sel=None, stream=None, out=None):
    """Fused octaves on device tensors.

    verts: float64 [V,3] cuda tensor.  elevations: float64 [V] cuda tensor
    the octaves are added to in place; or None, in which case the octave sum
    is written to `out` (allocated when None) without reading it.
    rough/strength: per-octave host arrays (see octave_scalars).
    """
    import torch
    from .opensimplex import _tables
    perm, pgi = _tables(perm, pgi)
    rough = np.ascontiguousarray(rough, dtype=np.float64)
    strength = np.ascontiguousarray(strength, dtype=np.float64)
    if rough.shape != strength.shape or rough.ndim != 1:
        raise ValueError("rough and strength must be 1-D of equal length")
    if verts.dtype != torch.float64 or verts.dim() != 2 or verts.shape[1] != 3:
        raise ValueError("verts must be a float64 [V,3] tensor")
    verts = verts.contiguous()
    V = verts.shape[0]
    accumulate = elevations is not None
    if elevations is None:
        elevations = out if out is not None else torch.empty(V, dtype=torch.float64,
                                                              device=verts.device)
    if elevations.dtype != torch.float64 or elevations.shape != (V,) or \
            not elevations.is_contiguous():
        raise ValueError("elevations must be a contiguous float64 [V] tensor")
    ctx = _lib.Context.get(verts.device)
    _lib.check(_lib.lib().nx_sample_octaves(
        ctx.handle, _lib.ptr(verts), V, _lib.ptr(perm), _lib.ptr(pgi), rough.size,
        _lib.ptr(rough), _lib.ptr(strength), float(radius), _lib.ptr(elevations),
        1 if accumulate else 0, _lib.ptr(dbg_noise), _lib.ptr(dbg_sel),
        _lib.stream_ptr(stream)))
    return elevations


def _to_device(a):
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)


def sample_noise(verts, perm, pgi, n_roughness=1, n_strength=0.2, radius=1):
    """terrain.py:12-28 -> new float64[V]: ((noise(v*n_roughness)+1)*0.5*n_strength)*radius."""
    verts = np.asarray(verts)
    dv = _to_device(verts)
    out = sample_octaves_device(dv, None, perm, pgi, [float(n_roughness)],
                                [float(n_strength)], float(radius))
    _lib.Context.get(dv.device).sync_check()
    return out.cpu().numpy()


def sample_octaves(verts, elevations, perm, pgi, n_octaves=1, n_init_roughness=1.5,
                   n_init_strength=0.4, n_roughness=2.0, n_persistence=0.5,
                   world_radius=1.0):
    """terrain.py:32-59.  Returns ``elevations`` (mutated in place when given)."""
    verts = np.asarray(verts)
    if elevations is None:
        elevations = np.zeros(len(verts), dtype=np.float64)
    if elevations.dtype != np.float64:
        raise TypeError("elevations must be float64")
    rough, strength = octave_scalars(n_octaves, n_init_roughness, n_init_strength,
                                     n_roughness, n_persistence, world_radius)
    dv = _to_device(verts)
    de = _to_device(elevations)
    sample_octaves_device(dv, de, perm, pgi, rough, strength, float(world_radius))
    _lib.Context.get(dv.device).sync_check()
    elevations[...] = de.cpu().numpy()
    return elevations


def make_bool_elevation_mask_device(height, mask_elevation, out=None, stream=None):
    """terrain.py:61-72 on a float64 device tensor: uint8 [n], 1 where
    height <= mask_elevation."""
    import torch
    if height.dtype != torch.float64 or not height.is_contiguous():
        raise ValueError("height must be a contiguous float64 tensor")
    if out is None:
        out = torch.empty(height.numel(), dtype=torch.uint8, device=height.device)
    ctx = _lib.Context.get(height.device)
    _lib.check(_lib.lib().nx_make_bool_elevation_mask(
        ctx.handle, _lib.ptr(height), height.numel(), float(mask_elevation), _lib.ptr(out),
        _lib.stream_ptr(stream)))
    return out


def make_bool_elevation_mask(height, mask_elevation):
    """terrain.py:61-72: a bool array, True where height <= mask_elevation."""
    import torch
    a = np.ascontiguousarray(height, dtype=np.float64).ravel()
    if a.size == 0:
        return np.zeros(0, dtype=np.bool_)
    dev = torch.device("cuda", torch.cuda.current_device())
    m = make_bool_elevation_mask_device(torch.from_numpy(a).to(dev), mask_elevation)
    _lib.Context.get(dev).sync_check()
    return m.cpu().numpy().astype(np.bool_)


def terrain_heights_device(verts, perm, pgi, rough, strength, radius=1.0, min_alt=-4000.0,
                           max_alt=8850.0, ocean_percent=55.0, stream=None, workspace=None,
                           out=None):
    """terrain.py:32-59 then nixis.py:330-364 in one call on device tensors:
    the octave sum at the vertices (from zero), its min/max reduced inside the
    noise kernel, then the height chain.  Returns (heights float64 [V], ocean
    mask uint8 [V], ocean level float64 [1]), all on the device; equal to
    sample_octaves_device followed by util.height_chain_device."""
    import torch
    from .opensimplex import _tables
    perm, pgi = _tables(perm, pgi)
    rough = np.ascontiguousarray(rough, dtype=np.float64)
    strength = np.ascontiguousarray(strength, dtype=np.float64)
    if rough.shape != strength.shape or rough.ndim != 1:
        raise ValueError("rough and strength must be 1-D of equal length")
    if verts.dtype != torch.float64 or verts.dim() != 2 or verts.shape[1] != 3:
        raise ValueError("verts must be a float64 [V,3] tensor")
    verts = verts.contiguous()
    V = verts.shape[0]
    L = _lib.lib()
    need = L.nx_height_chain_workspace_bytes(V)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=verts.device)
    h = out if out is not None else torch.empty(V, dtype=torch.float64, device=verts.device)
    mask = torch.empty(V, dtype=torch.uint8, device=verts.device)
    level = torch.empty(1, dtype=torch.float64, device=verts.device)
    ctx = _lib.Context.get(verts.device)
    _lib.check(L.nx_terrain_heights(
        ctx.handle, _lib.ptr(verts), V, _lib.ptr(perm), _lib.ptr(pgi), rough.size,
        _lib.ptr(rough), _lib.ptr(strength), float(radius), float(min_alt), float(max_alt),
        float(ocean_percent), _lib.ptr(h), _lib.ptr(mask), _lib.ptr(level),
        _lib.ptr(workspace), _lib.stream_ptr(stream)))
    return h, mask, level
