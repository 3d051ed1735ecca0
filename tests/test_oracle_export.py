"""Pin the oracle's equirectangular export on the CPU (util.py:80-90,
util.py:275-366, nixis.py:270-302).

* make_ll_arr: closed forms of the grid (north pole row first, the equator
  row and the 0-longitude column exactly on their planes, every point on the
  sphere, latitude and longitude recovered by asin / atan2).
* the k-nearest query: scipy's KDTree is the reference's own tree
  (util.py:285), importable here, so the oracle's brute force is checked
  against it directly, distances bit for bit.  Where distances tie exactly
  (mirror-image vertices seen from a pole pixel) the oracle orders the
  neighbours by index and scipy by its traversal; there the sets agree.
* make_gray_array: inverse-distance weights worked out by hand, including
  truncation toward zero (int(), not round or floor).
"""
import numpy as np
import pytest

from nixis_amd import util


def test_ll_grid_closed_forms(oracle):
    W, H, r = 16, 9, 2.5
    ll = oracle.make_ll_arr(W, H, r)
    assert ll.shape == (H, W, 3)
    assert (ll[0, :, 2] == r).all() and (ll[-1, :, 2] == -r).all()     # poles: rows 0 and H-1
    assert (ll[(H - 1) // 2, :, 2] == 0.0).all()                         # latitude 90 - 4*22.5 = 0
    assert (ll[:, W // 2, 1] == 0.0).all()                               # longitude -180 + 8*22.5 = 0
    assert (ll[1:-1, W // 2, 0] > 0).all()                               # ... on the +x side
    assert np.allclose(np.linalg.norm(ll, axis=2), r, rtol=1e-15, atol=0)
    lat = np.degrees(np.arcsin(ll[1:-1, :, 2] / r))
    lon = np.degrees(np.arctan2(ll[1:-1, :, 1], ll[1:-1, :, 0]))
    assert np.allclose(lat, (90 - 180 / (H - 1) * np.arange(1, H - 1))[:, None], atol=1e-12)
    want_lon = -180 + 360 / W * np.arange(W)
    got = (lon - want_lon[None, :] + 180) % 360 - 180                  # -180 and 180 are one line
    assert np.abs(got).max() < 1e-12


@pytest.mark.parametrize("k", [1, 3, 5])
def test_knn_equals_scipy_kdtree_random(oracle, k):
    from scipy.spatial import KDTree
    rng = np.random.default_rng(k)
    p = rng.standard_normal((4000, 3))
    q = rng.standard_normal((700, 3)) * 1.3
    d, n = oracle.knn(p, q, k)
    d2, n2 = KDTree(p, leafsize=10).query(q, k=k, workers=-1)
    d2 = d2.reshape(d.shape)
    n2 = n2.reshape(n.shape)
    assert np.array_equal(n, n2) and np.array_equal(d, d2)


def test_knn_equals_scipy_kdtree_on_pixels_of_an_icosphere(oracle):
    """The reference's use: pixel points of make_ll_arr against the mesh."""
    from scipy.spatial import KDTree
    pts, _ = util.create_mesh(24)
    r = 1.0
    ll = oracle.make_ll_arr(96, 48, r)
    d, n = oracle.knn(pts * r, ll, 3)
    d2, n2 = KDTree(pts * r, leafsize=10).query(ll, k=3, workers=-1)
    assert d.shape == (48, 96, 3) and n.dtype == np.int64
    assert np.array_equal(d, d2)
    # Mirror-image vertices can be exactly equidistant from a pixel (the pole
    # rows, symmetry planes of the icosphere): there the oracle orders by
    # index and scipy by its traversal, also across the k-th place.  Every
    # differing slot must hold two vertices at exactly that slot's distance.
    diff = n != n2
    assert diff.any(axis=2).mean() < 0.02
    P = pts * r
    for y, x, j in zip(*np.nonzero(diff)):
        for v in (n[y, x, j], n2[y, x, j]):
            e = ll[y, x] - P[v]
            s = 0.0
            for c in e:
                s += c * c
            assert np.sqrt(s) == d[y, x, j]


def test_gray_weights_by_hand(oracle):
    colors = np.array([10.0, 20.0, 40.0, 100.0, 0.0, -10.0, -20.0, -40.0])
    # equal distances: the plain mean, 70/3 = 23.33 -> 23
    px = oracle.make_gray_array(1, 1, [[1.0, 1.0, 1.0]], [[0, 1, 2]], colors)
    assert px.item() == 23
    # negative: -23.33 truncates toward zero to -23 (floor would give -24)
    px = oracle.make_gray_array(1, 1, [[1.0, 1.0, 1.0]], [[5, 6, 7]], colors)
    assert px.item() == -23
    # nearer is heavier: weights 1/d = 10, 5, 2.5 -> 100 * 2.5 / 17.5 = 14.29 -> 14
    # (weighting by d instead would give 57)
    px = oracle.make_gray_array(1, 1, [[0.1, 0.2, 0.4]], [[4, 4, 3]], colors)
    assert px.item() == 14
    # on a vertex: d = 0 gets weight 2 / 1e-5, the others 2 / 1.00001:
    # 100 * 200000 / 200003.99996 = 99.998 -> 99 (rounding would give 100)
    px = oracle.make_gray_array(1, 1, [[0.0, 1.0, 1.0]], [[3, 4, 4]], colors)
    assert px.item() == 99
    # layout: pixel p of a [2, 3] image uses row p of dists / nbrs
    d = np.ones((2, 3, 3))
    n = np.zeros((2, 3, 3), np.int64) + np.array([0, 1, 2])
    d[1, 2] = [0.0, 1.0, 1.0]
    n[1, 2] = [3, 4, 4]
    img = oracle.make_gray_array(3, 2, d, n, colors)
    assert img.shape == (2, 3) and img[1, 2] == 99 and (img.ravel()[:-1] == 23).all()
