"""Device mesh generation equals the host generator and table builders bit
for bit (util.py:17-50 stand-in, build_adjacency util.py:590-613,
sort_adjacency util.py:638-662), including the reference's append order
for a triangle list in an arbitrary order."""
import numpy as np
import pytest
import torch

from nixis_amd import _lib, util

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("d", [1, 2, 3, 8, 57, 400])
def test_device_icosphere_equals_host(d):
    pts, cells = util.create_mesh(d)
    dp, dc = util.create_mesh_device(d)
    assert (dp.cpu().numpy() == pts).all()
    assert (dc.cpu().numpy() == cells).all()


@pytest.mark.parametrize("d", [1, 5, 64])
def test_device_adjacency_equals_host(d):
    pts, cells = util.create_mesh(d)
    adj = util.build_adjacency(cells)
    dadj = util.build_adjacency_device(torch.from_numpy(cells).cuda(), pts.shape[0])
    assert (dadj.cpu().numpy() == adj).all()
    util.sort_adjacency(adj)
    sadj = util.sort_adjacency_device(dadj)
    assert (sadj.cpu().numpy() == adj).all()


def test_device_adjacency_keeps_triangle_order(oracle):
    """Rows follow the triangle order of the input, whatever it is."""
    pts, cells = util.create_mesh(9)
    perm = np.random.default_rng(3).permutation(cells.shape[0])
    shuffled = np.ascontiguousarray(cells[perm])
    ref = oracle.build_adjacency(shuffled, pts.shape[0])
    dadj = util.build_adjacency_device(torch.from_numpy(shuffled).cuda(), pts.shape[0])
    assert (dadj.cpu().numpy() == ref).all()
    oracle.sort_adjacency(ref)
    assert (util.sort_adjacency_device(dadj).cpu().numpy() == ref).all()


def test_device_adjacency_rejects_degree_over_six():
    cells = np.array([[0, k, k + 1] for k in range(1, 9)], dtype=np.int32)   # fan of 8
    with pytest.raises(_lib.NixisArgError):
        util.build_adjacency_device(torch.from_numpy(cells).cuda(), 12)


def test_device_mesh_at_scale():
    """d=2000 (40,000,002 vertices): points, triangles and sorted table."""
    d = 2000
    pts, nb = util.icosphere_with_neighbors(d)
    dp, dn = util.icosphere_with_neighbors_device(d)
    assert (dp.cpu().numpy() == pts).all()
    assert (dn.cpu().numpy() == nb).all()
