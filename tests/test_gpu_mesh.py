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


@pytest.mark.parametrize("d", [3, 40, 300])
def test_device_sfc_order_and_layout_equal_host(d):
    from nixis_amd import erosion
    pts, nb = util.icosphere_with_neighbors(d)
    dp, dn = util.icosphere_with_neighbors_device(d)
    order = util.sfc_order(pts)
    dorder = util.sfc_order_device(dp)
    assert (dorder.cpu().numpy() == order).all()
    g_host = erosion.ErosionGraph(pts, nb)
    g_dev = erosion.ErosionGraph.from_device(dp, dn)
    assert (g_dev.order == g_host.order).all()
    assert (g_dev.col.cpu().numpy() == g_host.col.cpu().numpy()).all()
    assert (g_dev.elen.cpu().numpy() == g_host.elen.cpu().numpy()).all()
    assert (g_dev.plan is None) == (g_host.plan is None)
    if g_host.plan is not None:
        assert (g_dev._plan_blob.cpu().numpy() == g_host._plan_blob.cpu().numpy()).all()


def test_mesh_cache_device_equals_host_path():
    from nixis_amd import batch
    job = batch.Job(d=48, seed=5, n_octaves=6, erosion_iters=17, rain=1.0)
    outs = []
    for on_device in (True, False):
        meshes = batch.MeshCache(on_device=on_device)
        run = batch.PlanetRun(job, meshes)
        batch.BatchRunner().run([run])
        outs.append(run.heights())
        assert (meshes.verts(48).cpu().numpy() ==
                util.create_points(48)[meshes.order(48)]).all()
    assert (outs[0] == outs[1]).all()
