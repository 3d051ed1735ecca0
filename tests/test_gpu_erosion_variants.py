"""GPU parity of erosion v1 and v4 (erosion.py:42-99, erosion.py:301-381)
against the FP64 oracle: bit for bit, whatever the numbering, tiling or
launch grouping (the kernels share the v3 gather and tile templates)."""
import numpy as np
import pytest
import torch

from nixis_amd import erosion, opensimplex as osx, terrain, util

pytestmark = pytest.mark.gpu


def _planet(d, seed=5):
    pts, nb = util.icosphere_with_neighbors(d)
    perm, pgi = osx.init(seed)
    h = terrain.sample_octaves(pts, None, perm, pgi, 8, 1.5, 0.4, 2.0, 0.5)
    return pts, nb, util.rescale(h, -9000, 9000)


def _run4(g, h, w, s, n, rain):
    st = erosion.ErosionState(g.V)
    st.load(*[torch.from_numpy(g.permute(a)).cuda() for a in (h, w, s)])
    st.run(g, n, rain, rule=4)
    out = [torch.empty(g.V, dtype=torch.float64, device="cuda") for _ in range(3)]
    st.store(*out)
    st.ctx.sync_check()
    return [g.unpermute(t.cpu().numpy()) for t in out]


def _oracle4(oracle, pts, nb, h, w, s, n, rain):
    h, w, s = h.copy(), w.copy(), s.copy()
    for _ in range(n):
        if rain:
            w += rain
        oracle.erosion_iteration4(pts, nb, h, w, s)
    return h, w, s


@pytest.mark.parametrize("tiled", [True, False])
def test_v4_steps_match_oracle(oracle, tiled):
    pts, nb, h = _planet(32)
    rng = np.random.default_rng(1)
    w = rng.uniform(-0.5, 1.0, h.size)          # negative water exercises the clamps
    s = rng.uniform(-0.01, 0.02, h.size)
    g = erosion.ErosionGraph(pts, nb, tiled=tiled)
    assert (g.plan is not None) == tiled
    for n, rain in ((1, 0.0), (7, 0.0), (40, 0.3 / 320)):
        print("v4 tiled", tiled, "steps", n, "rain", rain, flush=True)
        got = _run4(g, h, w, s, n, rain)
        ref = _oracle4(oracle, pts, nb, h, w, s, n, rain)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


def test_v4_drop_in_one_step(oracle):
    pts, nb, h = _planet(20)
    rng = np.random.default_rng(2)
    w = rng.uniform(-0.2, 0.5, h.size)
    s = rng.uniform(0.0, 0.01, h.size)
    a = [h.copy(), w.copy(), s.copy()]
    erosion.erosion_iteration4(pts, nb, *a)
    b = [h.copy(), w.copy(), s.copy()]
    oracle.erosion_iteration4(pts, nb, *b)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_v1_matches_oracle(oracle):
    pts, nb, h = _planet(40)
    for n in (1, 2, 25):
        got = erosion.erode_terrain1(pts, nb, h.copy(), num_iter=n)
        ref = h.copy()
        tmp = np.empty_like(ref)
        for _ in range(n):
            oracle.erosion_iteration1(nb, ref, tmp)
            ref[...] = tmp
        assert np.array_equal(got, ref)
    w = np.empty_like(h)
    assert erosion.erosion_iteration1(nb, h, w) is w
    ref = np.empty_like(h)
    oracle.erosion_iteration1(nb, h, ref)
    assert np.array_equal(w, ref)
    same = h.copy()
    erosion.erode_terrain1(pts, nb, same, num_iter=0)      # num_iter <= 0 runs one pass
    assert np.array_equal(same, erosion.erode_terrain1(pts, nb, h.copy(), num_iter=1))


def test_erode_terrain4_is_erode_terrain3(oracle):
    """The reference's erode_terrain4 calls erosion_iteration3 with 0.3/320
    of rain per pass."""
    pts, nb, h = _planet(24)
    a, b = h.copy(), h.copy()
    erosion.erode_terrain4(pts, nb, a, num_iter=30)
    oracle.erode_terrain3(pts, nb, b, num_iter=30, rain=0.3 / 320)
    assert np.array_equal(a, b)


def test_v4_full_size_tiled_equals_single_steps():
    """d=1000 (10,000,002 vertices): 20 steps two per launch over the tile
    plan equal 20 one-step launches."""
    pts, nb, h = _planet(1000)
    g = erosion.ErosionGraph(pts, nb)
    g1 = erosion.ErosionGraph(pts, nb, tiled=False)
    w = np.full(h.size, 0.01)
    s = np.zeros(h.size)
    a = _run4(g, h, w, s, 20, 0.3 / 320)
    b = _run4(g1, h, w, s, 20, 0.3 / 320)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.isfinite(a[0]).all() and (a[1] >= 0).all()
