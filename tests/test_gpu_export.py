"""GPU parity of the equirectangular export (util.py:275-455,
nixis.py:270-302) against the oracle and against scipy's KDTree.

* make_ll_arr: the per-row / per-column cosines and sines come from the
  host C library, as the reference's numba code calls it, so the pixel
  coordinates equal the oracle's bit for bit.
* the k-nearest query: exactly the oracle's (brute force, equal distances
  ordered by index), distances bit for bit; against scipy's KDTree (the
  reference's tree) the distances are equal and the indices differ only
  where two vertices are at exactly the same distance.
* make_gray_array: bit for bit with the oracle on the same query results.
"""
import os

import numpy as np
import pytest
import torch

from nixis_amd import _lib, util

pytestmark = pytest.mark.gpu


def _ties_only(d, n, d2, n2, pts, q):
    """Indices differ only where both vertices lie at exactly that distance."""
    assert np.array_equal(d, d2)
    bad = np.nonzero(n != n2)
    qq = q.reshape(-1, 3) if q.ndim > 2 else q
    flat = np.ravel_multi_index(bad[:-1], n.shape[:-1])
    for f, j in zip(flat, bad[-1]):
        for v in (n.reshape(-1, n.shape[-1])[f, j], n2.reshape(-1, n.shape[-1])[f, j]):
            e = qq[f] - pts[v]
            s = 0.0
            for c in e:
                s += c * c
            assert np.sqrt(s) == d.reshape(-1, d.shape[-1])[f, j]
    return len(flat)


@pytest.mark.parametrize("wh", [(16, 9), (96, 48), (2048, 1024), (1001, 333)])
def test_ll_arr_bit_exact(oracle, wh):
    w, h = wh
    print("ll", w, h, flush=True)
    got = util.make_ll_arr(w, h, 1.7)
    assert got.shape == (h, w, 3)
    assert np.array_equal(got, oracle.make_ll_arr(w, h, 1.7))


@pytest.mark.parametrize("k", [1, 3, 8])
def test_query_random_points_matches_oracle(oracle, k):
    rng = np.random.default_rng(40 + k)
    p = rng.standard_normal((20000, 3))
    q = np.concatenate([rng.standard_normal((3000, 3)) * 1.5,
                        p[:50] + 1e-12,                      # on top of a point
                        np.full((3, 3), 25.0)])              # far outside the grid
    print("random k", k, flush=True)
    grid = util.KnnGrid(p)
    d, n = grid.query(q, k=k)
    rd, rn = oracle.knn(p, q, k)
    if k == 1:
        rd, rn = rd[..., 0], rn[..., 0]
    assert d.shape == rd.shape and n.dtype == np.int64
    assert np.array_equal(d, rd) and np.array_equal(n, rn)


def test_query_fewer_points_than_k(oracle):
    p = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    q = np.array([[0.2, 0, 0], [5.0, 1.0, -2.0]])
    d, n = util.KnnGrid(p).query(q, k=3)
    rd, rn = oracle.knn(p, q, 3)
    assert n.tolist() == [[0, 1, 2], [1, 0, 2]] and (d[:, 2] == np.inf).all()  # scipy: n, inf
    assert np.array_equal(d, rd) and np.array_equal(n, rn)


def test_query_pixels_matches_oracle_and_scipy(oracle):
    from scipy.spatial import KDTree
    pts, _ = util.create_mesh(64)
    r = 2.0
    P = pts * r
    grid = util.KnnGrid(P)
    d, n = grid.query_pixels_device(256, 128, r, 3)
    d, n = d.cpu().numpy(), n.cpu().numpy()
    ll = oracle.make_ll_arr(256, 128, r)
    rd, rn = oracle.knn(P, ll, 3)
    assert np.array_equal(d, rd) and np.array_equal(n, rn)
    sd, sn = KDTree(P, leafsize=10).query(ll, k=3, workers=-1)
    ties = _ties_only(d, n, sd, sn, P, ll)
    print(f"pixels 256x128 on d=64: {ties} slots ordered differently from scipy at exact ties",
          flush=True)


def test_query_pixels_full_image_against_scipy():
    """2048 x 1024 (the reference's default image) on a d=500 mesh
    (2,500,002 vertices) against scipy's KDTree on the host."""
    from scipy.spatial import KDTree
    pts = util.create_points(500)
    grid = util.KnnGrid(pts)
    d, n = grid.query_pixels_device(2048, 1024, 1.0, 3)
    d, n = d.cpu().numpy(), n.cpu().numpy()
    ll = util.make_ll_arr(2048, 1024, 1.0)
    sd, sn = KDTree(pts, leafsize=10).query(ll, k=3, workers=-1)
    ties = _ties_only(d, n, sd, sn, pts, ll)
    print(f"2048x1024 on d=500: {ties} slots at exact ties", flush=True)
    assert ties < 0.001 * n.size


def test_gray_array_bit_exact(oracle):
    rng = np.random.default_rng(3)
    pts, _ = util.create_mesh(40)
    grid = util.KnnGrid(pts)
    d, n = grid.query_pixels_device(300, 150, 1.0, 3)
    d, n = d.cpu().numpy(), n.cpu().numpy()
    for colors in (rng.uniform(0, 255, pts.shape[0]),
                   rng.uniform(-300, 300, pts.shape[0]),
                   rng.integers(0, 65535, pts.shape[0]).astype(np.uint16)):
        got = util.make_gray_array(300, 150, d, n, colors)
        ref = oracle.make_gray_array(300, 150, d, n, colors).astype(colors.dtype)
        assert got.dtype == colors.dtype and np.array_equal(got, ref)
    rgb = util.make_rgb_array(300, 150, d, n, colors)
    assert rgb.shape == (150, 300, 3) and (rgb == got[:, :, None]).all()


def test_gray_rejects_neighbours_outside_colors():
    d = torch.ones((4, 3), dtype=torch.float64, device="cuda")
    n = torch.zeros((4, 3), dtype=torch.int64, device="cuda")
    n[2, 1] = 10
    c = torch.ones(10, dtype=torch.float64, device="cuda")
    util.make_gray_array_device(d, n, c)
    with pytest.raises(_lib.NixisArgError):
        _lib.Context.get(d.device).sync_check()


def test_export_pipeline_writes_the_oracle_image(oracle, tmp_path):
    """heights -> build_image_data (rescale to 0..255, gray) -> save_image:
    the PNG read back equals the oracle's pixels."""
    from PIL import Image
    pts, _ = util.create_mesh(50)
    rng = np.random.default_rng(9)
    h = rng.uniform(-4000, 8850, pts.shape[0])
    ocean = h <= 0
    grid = util.build_KDTree(pts)
    qd = grid.query(util.make_ll_arr(400, 200, 1.0), k=3, workers=-1)
    data = util.build_image_data({"height": [h, "gray"], "ocean": [ocean, "gray"]}, 400, 200, qd)
    util.save_image(data, str(tmp_path), "w")
    img = np.asarray(Image.open(os.path.join(tmp_path, "w_height.png")))
    ref = oracle.make_gray_array(400, 200, qd[0], qd[1],
                                 oracle.rescale_ex(h, 0, 255)).astype(np.uint8)
    assert img.dtype == np.uint8 and np.array_equal(img, ref)
    m = np.asarray(Image.open(os.path.join(tmp_path, "w_ocean.png")))
    ref_m = oracle.make_gray_array(400, 200, qd[0], qd[1],
                                   oracle.rescale_ex(ocean.astype(np.float64), 0, 255))
    assert np.array_equal(m, ref_m.astype(np.float64).astype(np.uint8))
