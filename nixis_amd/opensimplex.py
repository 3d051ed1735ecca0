"""Drop-in for the reference's opensimplex.py (3D path only).

Same names, signatures and return types as opensimplex.py:90-107 (init),
257-263 (noisearr3d) and 266-759 (noise3d).  ``init`` runs on the host in
the C library (2 KB of tables); the noise functions evaluate on the GPU:
lattice selection bit-exact with the reference (FP64), gradient/falloff
arithmetic in FP32 (|error| <= 1e-5).  The 2D/4D functions of the reference
are not on the terrain path and are not provided.
"""
import numpy as np

from . import _lib

# opensimplex.py:26-27,32
STRETCH_CONSTANT_3D = -1.0 / 6
SQUISH_CONSTANT_3D = 1.0 / 3
NORM_CONSTANT_3D = 103
DEFAULT_SEED = 0

# opensimplex.py:51-60
GRADIENTS_3D = (
    -11, 4, 4, -4, 11, 4, -4, 4, 11,
    11, 4, 4, 4, 11, 4, 4, 4, 11,
    -11, -4, 4, -4, -11, 4, -4, -4, 11,
    11, -4, 4, 4, -11, 4, 4, -4, 11,
    -11, 4, -4, -4, 11, -4, -4, 4, -11,
    11, 4, -4, 4, 11, -4, 4, 4, -11,
    -11, -4, -4, -4, -11, -4, -4, -4, -11,
    11, -4, -4, 4, -11, -4, 4, -4, -11,
)

_INT64_MIN = -(1 << 63)
_INT64_MAX = (1 << 63) - 1


def init(seed=DEFAULT_SEED):
    """opensimplex.py:90-107 -> (perm int32[256], perm_grad_index_3D int32[256]).

    The seed is an int64 (numba types a Python int argument as int64; larger
    magnitudes are rejected rather than reproducing numba's uint64/float
    promotion).
    """
    seed = int(seed)
    if not (_INT64_MIN <= seed <= _INT64_MAX):
        raise OverflowError(f"seed {seed} does not fit int64")
    perm = np.zeros(256, dtype=np.int32)
    pgi = np.zeros(256, dtype=np.int32)
    _lib.check(_lib.lib().nx_init_tables(seed, _lib.ptr(perm), _lib.ptr(pgi)))
    return perm, pgi


def _tables(perm, pgi):
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    pgi = np.ascontiguousarray(pgi, dtype=np.int32)
    if perm.shape != (256,) or pgi.shape != (256,):
        raise ValueError("perm and perm_grad_index_3D must have 256 entries")
    return perm, pgi


def noisearr3d_device(x, y, z, perm, perm_grad_index_3D, stream=None):
    """noisearr3d on device tensors (float64, same shape) -> device float64."""
    import torch
    perm, pgi = _tables(perm, perm_grad_index_3D)
    x, y, z = (t.contiguous().to(torch.float64) for t in (x, y, z))
    out = torch.empty_like(x)
    ctx = _lib.Context.get(x.device)
    _lib.check(_lib.lib().nx_noise3d(ctx.handle, _lib.ptr(x), _lib.ptr(y), _lib.ptr(z), x.numel(),
                                     _lib.ptr(perm), _lib.ptr(pgi), _lib.ptr(out),
                                     _lib.stream_ptr(stream)))
    return out


def noisearr3d(x, y, z, perm, perm_grad_index_3D):
    """opensimplex.py:257-263: noise at every (x[i], y[i], z[i]) -> float64[N]."""
    import torch
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.ascontiguousarray(y, dtype=np.float64).ravel()
    z = np.ascontiguousarray(z, dtype=np.float64).ravel()
    if not (x.size == y.size == z.size):
        raise ValueError("x, y, z must have the same size")
    if x.size == 0:
        return np.zeros(0, dtype=np.float64)
    dev = torch.device("cuda", torch.cuda.current_device())
    tx, ty, tz = (torch.from_numpy(a).to(dev) for a in (x, y, z))
    out = noisearr3d_device(tx, ty, tz, perm, perm_grad_index_3D)
    _lib.Context.get(dev).sync_check()
    return out.cpu().numpy()


def noise3d(x, y, z, perm, perm_grad_index_3D):
    """opensimplex.py:266-759: noise at one point -> float."""
    return float(noisearr3d(np.array([x]), np.array([y]), np.array([z]), perm,
                            perm_grad_index_3D)[0])
