"""nx_terrain_heights_noalloc (static device scratch, no allocation) against
nx_terrain_heights (caller workspace), through the C ABI.

Two calls in a row on one GPU, the second on a smaller planet, each compared
bit for bit with the allocating entry point on the same inputs: heights,
ocean mask and ocean level.  The second call reuses the static scratch the
first left behind, so anything it failed to reset would show here.  The
file is named to run last in the GPU suite, after every other test has used
the context.
"""
import ctypes

import numpy as np
import pytest
import torch

from nixis_amd import _lib, opensimplex as osx, terrain, util

pytestmark = pytest.mark.gpu


def _call(entry, verts, perm, pgi, rough, strength, ws=None):
    V = verts.shape[0]
    h = torch.full((V,), np.nan, dtype=torch.float64, device=verts.device)
    mask = torch.full((V,), 7, dtype=torch.uint8, device=verts.device)
    level = torch.full((1,), np.nan, dtype=torch.float64, device=verts.device)
    ctx = _lib.Context.get(verts.device)
    L = _lib.lib()
    args = [ctx.handle, _lib.ptr(verts), V, _lib.ptr(perm), _lib.ptr(pgi), rough.size,
            _lib.ptr(rough), _lib.ptr(strength), 1.0, -4000.0, 8850.0, 55.0, _lib.ptr(h),
            _lib.ptr(mask), _lib.ptr(level)]
    if entry == "nx_terrain_heights":
        args.append(_lib.ptr(ws))
    args.append(_lib.stream_ptr())
    _lib.check(getattr(L, entry)(*args))
    ctx.sync_check()
    return h.cpu().numpy(), mask.cpu().numpy(), level.cpu().numpy()


@pytest.mark.parametrize("sizes", [((200, 1234, 8), (40, 7, 12))])
def test_noalloc_equals_workspace_entry_point(sizes):
    L = _lib.lib()
    for d, seed, n_oct in sizes:
        pts, _ = util.create_mesh_device(d, cells=False)
        perm, pgi = osx.init(seed)
        perm, pgi = np.ascontiguousarray(perm, np.int32), np.ascontiguousarray(pgi, np.int32)
        rough, strength = terrain.octave_scalars(n_oct, 1.5, 0.4, 2.0, 0.5, 1.0)
        rough = np.ascontiguousarray(rough, np.float64)
        strength = np.ascontiguousarray(strength, np.float64)
        V = pts.shape[0]
        print(f"d={d} V={V} seed={seed} octaves={n_oct}", flush=True)
        ws = torch.empty(int(L.nx_height_chain_workspace_bytes(V)), dtype=torch.uint8,
                         device=pts.device)
        want = _call("nx_terrain_heights", pts, perm, pgi, rough, strength, ws)
        got = _call("nx_terrain_heights_noalloc", pts, perm, pgi, rough, strength)
        assert np.array_equal(got[0].view(np.int64), want[0].view(np.int64)), "heights"
        assert np.array_equal(got[1], want[1]), "ocean mask"
        assert np.array_equal(got[2].view(np.int64), want[2].view(np.int64)), "ocean level"
        assert np.isfinite(got[0]).all() and set(np.unique(got[1])) <= {0, 1}
        assert got[0].min() == -4000.0 and got[0].max() == 8850.0


def test_noalloc_rejects_a_null_context():
    L = _lib.lib()
    with pytest.raises(_lib.NixisArgError):
        _lib.check(L.nx_terrain_heights_noalloc(None, None, 0, None, None, 0, None, None, 1.0,
                                                -1.0, 1.0, 50.0, None, None, None, None))
