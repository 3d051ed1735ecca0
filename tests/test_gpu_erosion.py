"""GPU parity of the erosion kernels against the FP64 oracle.

The kernels compute each step in FP64 in the reference's operation order
(erosion.py:209-267), so heights, water and sediment must equal the
oracle's bit for bit, whatever the vertex numbering, tiling or launch
grouping.  The stated bar, max |h_gpu - h_fp64| <= 1e-4 of the elevation
range after N iterations, is printed beside each check.
"""
import numpy as np
import pytest
import torch

from nixis_amd import _lib, erosion, opensimplex as osx, terrain, util

pytestmark = pytest.mark.gpu


def _planet(d, seed=99, n_oct=12):
    pts, nb = util.icosphere_with_neighbors(d)
    perm, pgi = osx.init(seed)
    h = terrain.sample_octaves(pts, None, perm, pgi, n_oct, 1.5, 0.4, 2.0, 0.5)
    h = util.rescale(h, -9000, 9000)
    return pts, nb, h


def _gpu_erode(pts, nb, h, iters, rain, order=None):
    g = erosion.ErosionGraph(pts, nb, order=order)
    st = erosion.ErosionState(g.V)
    st.load(torch.from_numpy(g.permute(h)).cuda())
    st.run(g, iters, rain)
    out = [torch.empty(g.V, dtype=torch.float64, device="cuda") for _ in range(3)]
    st.store(*out)
    st.ctx.sync_check()
    return [g.unpermute(t.cpu().numpy()) for t in out], st


def _rel(a, ref):
    return np.abs(a - ref).max() / max(ref.max() - ref.min(), 1e-300)


@pytest.mark.parametrize("d,iters,rain", [(16, 500, 1.0), (32, 300, 1.0),
                                          (32, 200, 0.3 / 320)])
def test_erosion_long_run_matches_fp64(oracle, d, iters, rain):
    pts, nb, h = _planet(d)
    (hg, wg, sg), _ = _gpu_erode(pts, nb, h, iters, rain)
    href = h.copy()
    w, s = oracle.erode_terrain3(pts, nb, href, iters, rain, return_fields=True)
    eh, ew, es = _rel(hg, href), _rel(wg, w), _rel(sg, s)
    print(f"\nd={d} iters={iters} rain={rain:g}: rel err h {eh:.2e} w {ew:.2e} s {es:.2e}")
    assert (hg == href).all() and (wg == w).all() and (sg == s).all()


def test_erosion_config_plain_small(oracle):
    """d=200, noise 8 oct seed 1234 rescaled to [-9000, 9000], 50 iterations
    of (rain 0.3/320, erosion_iteration3)."""
    pts, nb = util.icosphere_with_neighbors(200)
    perm, pgi = osx.init(1234)
    h = util.rescale(terrain.sample_octaves(pts, None, perm, pgi, 8, 1.5, 0.4, 2.0, 0.5),
                     -9000, 9000)
    hg = h.copy()
    erosion.erode_terrain3(pts, nb, hg, 50)
    href = h.copy()
    oracle.erode_terrain3(pts, nb, href, 50)
    assert (hg == href).all(), _rel(hg, href)


def test_erosion_sfc_order_same_result(oracle):
    pts, nb, h = _planet(24)
    order = util.sfc_order(pts)
    (h1, _, _), _ = _gpu_erode(pts, nb, h, 120, 1.0)
    (h2, _, _), _ = _gpu_erode(pts, nb, h, 120, 1.0, order=order)
    href = h.copy()
    oracle.erode_terrain3(pts, nb, href, 120, 1.0)
    assert (h1 == href).all() and (h2 == href).all()


def test_erosion_iteration3_single_step(oracle):
    rng = np.random.default_rng(4)
    pts, nb = util.icosphere_with_neighbors(12)
    V = pts.shape[0]
    h = rng.normal(size=V) * 100
    w = rng.uniform(0, 2, size=V)
    s = rng.uniform(0, 0.01, size=V)
    a = [h.copy(), w.copy(), s.copy()]
    b = [h.copy(), w.copy(), s.copy()]
    erosion.erosion_iteration3(pts, nb, *a)
    oracle.erosion_iteration3(pts, nb, *b)
    for x, y in zip(a, b):
        assert (x == y).all()


def test_erosion_flat_field_closed_form():
    """All heights equal: every neighbour ties, so s stays 0, h stays put and
    water follows w_k = r (2-e) ((2-e)^k - 1) / (1-e) (erosion.py:182,264)."""
    pts, nb = util.icosphere_with_neighbors(8)
    h = np.full(pts.shape[0], 123.0)
    k, r, e = 40, 0.3 / 320, 0.1 / 320
    (hg, wg, sg), _ = _gpu_erode(pts, nb, h, k, r)
    expect = r * (2 - e) * ((2 - e) ** k - 1) / (1 - e)
    assert (hg == 123.0).all() and (sg == 0).all()
    assert np.abs(wg / expect - 1).max() <= 1e-12


def test_erosion_mass_per_vertex_conserved():
    """h + s is conserved per vertex by every step (erosion.py:262-267); the
    FP64 mass check sees it to within FP64 rounding."""
    pts, nb, h = _planet(16)
    g = erosion.ErosionGraph(pts, nb)
    st = erosion.ErosionState(g.V)
    st.load(torch.from_numpy(h).cuda())
    m0 = st.mass().cpu().numpy()
    st.run(g, 200, 1.0)
    m1 = st.mass().cpu().numpy()
    hh, ss = (torch.empty(g.V, dtype=torch.float64, device="cuda") for _ in range(2))
    st.store(hh, None, ss)
    st.ctx.sync_check()
    hs = hh.cpu().numpy() + ss.cpu().numpy()
    scale = np.abs(hh.cpu().numpy()).max()
    assert np.abs(hs - h).max() <= 1e-12 * scale
    assert abs(m1[0] - m0[0]) <= 1e-12 * m1[1]


def test_erosion_chunked_runs_equal_one_run():
    """Running 70 + 50 iterations equals 120 in one call (state, exponent and
    pending rain carry over)."""
    pts, nb, h = _planet(12)
    g = erosion.ErosionGraph(pts, nb)
    outs = []
    for chunks in ([120], [70, 50]):
        st = erosion.ErosionState(g.V)
        st.load(torch.from_numpy(h).cuda())
        for c in chunks:
            st.run(g, c, 1.0)
        t = torch.empty(g.V, dtype=torch.float64, device="cuda")
        st.store(t)
        outs.append(t.cpu().numpy())
    assert (outs[0] == outs[1]).all()


def test_erosion_at_scale_short(oracle):
    """d=1000 (10,000,002 vertices), 12-octave seed-99 terrain, rain 1.0: 20
    iterations against the FP64 oracle at full size; the 500-iteration run
    is checked through size-independent properties in the bench."""
    pts, nb, h = _planet(1000)
    (hg, _, _), _ = _gpu_erode(pts, nb, h, 20, 1.0)
    href = h.copy()
    oracle.erode_terrain3(pts, nb, href, 20, 1.0)
    assert (hg == href).all(), _rel(hg, href)


def test_erosion_rejects_bad_args():
    pts, nb, h = _planet(4, n_oct=2)
    g = erosion.ErosionGraph(pts, nb)
    st = erosion.ErosionState(g.V)
    with pytest.raises(_lib.NixisArgError):
        st.run(g, -1)
    with pytest.raises(ValueError):
        erosion.ErosionGraph(pts, nb[:, :5])


def test_erosion_graph_replays_match_eager():
    """Captured chunks (16 steps) replayed several times, with eager work and
    synchronisations in between, equal the eager launches bit for bit."""
    pts, nb, h = _planet(20)
    g = erosion.ErosionGraph(pts, nb)
    ctx = _lib.Context.get(g.device)
    dh = torch.from_numpy(h).cuda()

    def run(chunk):
        ctx.set_graph_chunk(chunk)
        st = erosion.ErosionState(g.V)
        st.load(dh)
        st.run(g, 100, 1.0)
        out = torch.empty(g.V, dtype=torch.float64, device="cuda")
        st.store(out)
        ctx.sync_check()
        return out.cpu().numpy(), st

    try:
        eager, _ = run(0)
        ctx.set_graph_chunk(16)
        st = erosion.ErosionState(g.V)
        for rep in range(3):
            st.load(dh)
            st.run(g, 100, 1.0)          # 6 replays of the 16-step graph + 4
            out = torch.empty(g.V, dtype=torch.float64, device="cuda")
            st.store(out)
            # eager work between replays
            util.rescale_device(dh, -1, 1)
            torch.cuda.synchronize()
            assert (out.cpu().numpy() == eager).all(), f"replay {rep} differs"
    finally:
        ctx.set_graph_chunk(32)


def _run_graph(g, h, iters, rain, chunk=32):
    ctx = _lib.Context.get(g.device)
    ctx.set_graph_chunk(chunk)
    try:
        st = erosion.ErosionState(g.V)
        st.load(torch.from_numpy(g.permute(h)).cuda())
        st.run(g, iters, rain)
        out = [torch.empty(g.V, dtype=torch.float64, device="cuda") for _ in range(3)]
        st.store(*out)
        ctx.sync_check()
    finally:
        ctx.set_graph_chunk(32)
    return [g.unpermute(t.cpu().numpy()) for t in out]


@pytest.mark.parametrize("steps", [2, 4])
@pytest.mark.parametrize("d,iters,chunk", [(20, 101, 32), (24, 64, 0), (20, 75, 7), (2, 9, 32),
                                          (8, 10, 0), (16, 7, 6)])
def test_pair_launches_match_single_steps(oracle, d, iters, chunk, steps):
    """Two or four steps per launch over the tile plan (a four-step run's
    last two or three steps in one tail launch, a last single step alone;
    graph chunks of 7 and 6 end every chunk that way) give the one-step-per-
    launch results and the FP64 oracle's, bit for bit."""
    pts, nb, h = _planet(d)
    gt = erosion.ErosionGraph(pts, nb, steps=steps)
    g1 = erosion.ErosionGraph(pts, nb, tiled=False)
    assert gt.plan is not None and gt.plan.ntiles >= 1 and g1.plan is None
    assert gt.steps == steps and gt.plan.steps == steps
    a = _run_graph(gt, h, iters, 1.0, chunk)
    b = _run_graph(g1, h, iters, 1.0, chunk)
    href = h.copy()
    w, s = oracle.erode_terrain3(pts, nb, href, iters, 1.0, return_fields=True)
    for x, y, r in zip(a, b, (href, w, s)):
        assert (x == y).all() and (x == r).all()
    print(f"\nd={d} iters={iters} chunk={chunk}: tiles {gt.plan.ntiles}, "
          f"pair vs single {_rel(a[0], b[0]):.1e}, vs fp64 {_rel(a[0], href):.1e}")


@pytest.mark.parametrize("variant", range(7))
def test_pair_kernel_variants_match_oracle(oracle, variant):
    """Every two-step kernel variant (block size, neighbour data kept in
    registers or read again) gives the FP64 oracle's fields bit for bit, from
    one tile to over a thousand (odd step counts end with a single step)."""
    for d, iters, chunk in ((2, 5, 0), (24, 12, 0), (60, 7, 256), (160, 4, 256), (200, 3, 0)):
        pts, nb, h = _planet(d)
        g = erosion.ErosionGraph(pts, nb, chunk=chunk, variant=variant, steps=2)
        assert g.plan is not None and g.plan.steps == 2
        print(f"variant {variant} d={d} iters={iters} tiles {g.plan.ntiles}", flush=True)
        got = _run_graph(g, h, iters, 1.0)
        href = h.copy()
        w, s = oracle.erode_terrain3(pts, nb, href, iters, 1.0, return_fields=True)
        for x, r in zip(got, (href, w, s)):
            assert np.array_equal(x.view(np.int64), r.view(np.int64))


def test_graph_without_room_for_four_rings_takes_a_two_step_plan(oracle, monkeypatch):
    """When the planner refuses four rings (NX_ERANGE, forced here for every
    four-step request) the graph cuts a two-step plan instead, and erosion
    on it equals the oracle bit for bit."""
    pts, nb, h = _planet(24)
    real = _lib.lib()

    class FourStepsRefused:
        def __getattr__(self, name):
            return getattr(real, name)

        @staticmethod
        def nx_tile_plan_build_chunk(col, elen, V, steps, chunk, out, cap, plan):
            if steps == 4:
                return _lib.NX_ERANGE
            return real.nx_tile_plan_build_chunk(col, elen, V, steps, chunk, out, cap, plan)

    monkeypatch.setattr(_lib, "lib", lambda: FourStepsRefused())
    g = erosion.ErosionGraph(pts, nb)
    monkeypatch.undo()
    assert g.steps == 2 and g.plan is not None and g.plan.steps == 2
    got = _run_graph(g, h, 9, 1.0)
    href = h.copy()
    w, s = oracle.erode_terrain3(pts, nb, href, 9, 1.0, return_fields=True)
    for x, r in zip(got, (href, w, s)):
        assert np.array_equal(x, r)


@pytest.mark.parametrize("rule", [3, 4])
def test_four_step_plans_match_oracle(oracle, rule):
    """Four steps per launch, a remainder of two or three steps in one tail
    launch and of one step alone: 1..9 steps equal the oracle bit for bit,
    for every kernel variant."""
    pts, nb, h = _planet(32)
    V = h.size
    rng = np.random.default_rng(rule)
    w0 = rng.uniform(0.0, 0.02, V)
    s0 = rng.uniform(0.0, 0.001, V)
    g = erosion.ErosionGraph(pts, nb, steps=4)
    assert g.plan is not None and g.plan.steps == 4
    ref_step = oracle.erosion_iteration3 if rule == 3 else oracle.erosion_iteration4
    for variant in range(5):
        g_ctx = _lib.Context.get(g.device)
        g_ctx.set_multi_kernel(variant)
        for n in (1, 2, 3, 4, 5, 6, 7, 9):
            print("rule", rule, "variant", variant, "steps", n, flush=True)
            st = erosion.ErosionState(V)
            st.load(*[torch.from_numpy(g.permute(a)).cuda() for a in (h, w0, s0)])
            st.run(g, n, 0.0, rule=rule)
            out = [torch.empty(V, dtype=torch.float64, device="cuda") for _ in range(3)]
            st.store(*out)
            st.ctx.sync_check()
            hr, wr, sr = h.copy(), w0.copy(), s0.copy()
            for _ in range(n):
                ref_step(pts, nb, hr, wr, sr)
            for t, r in zip(out, (hr, wr, sr)):
                assert np.array_equal(g.unpermute(t.cpu().numpy()), r)
    _lib.Context.get(g.device).set_multi_kernel(0)


def test_four_step_launches_equal_two_step_and_single_at_scale():
    """d=1000 (10,000,002 vertices), 23 steps with rain (five four-step
    launches and a three-step tail): four-step plans (the default kernel,
    first-step data in registers, and variant 1, slot tables in LDS), the
    two-step plan and one step per launch give the same fields bit for
    bit."""
    pts, nb, h = _planet(1000, n_oct=8)
    outs = []
    ctx = None
    try:
        for kw, mv in (({"steps": 4}, 0), ({"steps": 4}, 1), ({"steps": 2}, 0), ({"tiled": False}, 0)):
            g = erosion.ErosionGraph(pts, nb, **kw)
            ctx = _lib.Context.get(g.device)
            ctx.set_multi_kernel(mv)
            st = erosion.ErosionState(g.V)
            st.load(torch.from_numpy(g.permute(h)).cuda())
            st.run(g, 23, 0.3 / 320)
            o = [torch.empty(g.V, dtype=torch.float64, device="cuda") for _ in range(3)]
            st.store(*o)
            st.ctx.sync_check()
            outs.append([g.unpermute(t.cpu().numpy()) for t in o])
            print("plan", kw, "four-step variant", mv, "done", flush=True)
    finally:
        if ctx is not None:
            ctx.set_multi_kernel(0)
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)


def test_graph_cache_equal_v_meshes_back_to_back(oracle):
    """Two meshes with the same vertex count but different tile plans (the
    same icosphere with jittered coordinates, so other edge lengths and
    another space-filling-curve order), eroded one after the other with
    captured graphs, the first mesh's graph and state freed before the next
    is built: the second run must not replay a graph captured for the first
    plan when its buffers land at the same addresses.  Every run equals the
    oracle bit for bit."""
    pts, nb, h = _planet(24)
    rng = np.random.default_rng(5)
    jit = pts + rng.normal(scale=0.02, size=pts.shape)
    jit /= np.linalg.norm(jit, axis=1, keepdims=True)
    seen = []
    for k, p in enumerate((pts, jit, pts, jit)):
        g = erosion.ErosionGraph(p, nb)
        assert g.plan is not None
        st = erosion.ErosionState(g.V)
        st.load(torch.from_numpy(g.permute(h)).cuda())
        st.run(g, 70, 1.0)                 # 32 + 32 + 6 steps, each a cached graph
        out = torch.empty(g.V, dtype=torch.float64, device="cuda")
        st.store(out)
        st.ctx.sync_check()
        got = g.unpermute(out.cpu().numpy())
        seen.append((st.buf.data_ptr(), g.plan.blob, g.plan.ntiles))
        print(f"\nrun {k}: state {seen[-1][0]:#x} plan {seen[-1][1]:#x} tiles {seen[-1][2]}",
              flush=True)
        del g, st, out
        torch.cuda.synchronize()
        href = h.copy()
        oracle.erode_terrain3(p, nb, href, 70, 1.0)
        assert (got == href).all(), f"run {k} differs from the oracle"


def test_reference_main_flow_at_world_radius(oracle):
    """The reference's own run (nixis.py:247-249, 308-364, 410): points
    scaled to the world radius 6378100 m, 7 octaves with roughness 1.5 x 2.5
    per octave and strength 0.4 x 0.5, the height chain to [-4000, 8850] m
    with the ocean at 55%, then erode_terrain3(points, neighbors, height,
    num_iter=11).  Noise within its per-octave tolerance of the oracle, the
    chain's mask and ocean level exact, erosion bit for bit from the same
    heights, through the drop-in erode_terrain3 on the scaled points."""
    R = 6378100.0
    d = 320
    pts, nb = util.icosphere_with_neighbors(d)
    pts = pts * R
    perm, pgi = osx.init(1234)
    n_oct, r0, s0, lac, pers = 7, 1.5, 0.4, 2.5, 0.5
    h = terrain.sample_octaves(pts, None, perm, pgi, n_oct, r0, s0, lac, pers, R)
    raw = oracle.sample_octaves(pts, None, perm, pgi, n_oct, r0, s0, lac, pers, R)
    # octave o adds ((n + 1) / 2 * (a_o / R)) * R: the radius cancels
    bound = 0.5 * sum(s0 * pers ** k for k in range(n_oct)) * 1e-5
    err = np.abs(h - raw).max()
    print(f"\nnoise at R={R}: max |elev - fp64| {err:.3e} m (bound {bound:.3e})")
    assert err <= bound
    hd = torch.from_numpy(h.copy()).cuda()
    hd, mask, level = util.height_chain_device(hd, -4000.0, 8850.0, 55.0)
    _lib.Context.get(hd.device).sync_check()
    ref_h, ref_mask, ref_level = oracle.height_chain(h, -4000.0, 8850.0, 55.0)
    assert level.item() == ref_level
    assert (mask.cpu().numpy().astype(bool) == ref_mask).all()
    hc = hd.cpu().numpy()
    assert np.abs(hc - ref_h).max() <= 64 * 2.0 ** -52 * 12850.0
    hg = hc.copy()
    erosion.erode_terrain3(pts, nb, hg, num_iter=11)
    href = hc.copy()
    oracle.erode_terrain3(pts, nb, href, 11)
    print(f"erosion at R={R}, 11 passes: max |dh| {np.abs(hg - hc).max():.3e} m")
    assert (hg == href).all(), _rel(hg, href)


def _same_bits_or_both_nan(a, b):
    """Equal bit patterns, except that NaNs only need to be NaN in both:
    the sign and payload of a NaN made by an invalid operation (inf - inf)
    are the hardware's choice (x86 makes 0xfff8..., the GPU 0x7ff8...)."""
    na, nb_ = np.isnan(a), np.isnan(b)
    return (na == nb_).all() and (a[~na].view(np.int64) == b[~nb_].view(np.int64)).all()


def test_erosion_into_overflow_matches_oracle(oracle):
    """The unit-radius planet with rain 1.0 grows about 2x per step
    (erosion.py:214-215,262-267 start each sum at the vertex's own value)
    until FP64 overflows near step 1002; inf - inf then makes NaN that
    spreads over the mesh.  Heights, water and sediment equal the oracle's
    at steps 1000..1005 and 1040 (infinities with their signs, finite values
    bit for bit, NaNs where the oracle has NaNs)."""
    pts, nb, h = _planet(16)
    g = erosion.ErosionGraph(pts, nb)
    st = erosion.ErosionState(g.V)
    st.load(torch.from_numpy(g.permute(h)).cuda())
    hr = h.copy()
    wr, sr = oracle.erode_terrain3(pts, nb, hr, 1000, 1.0, return_fields=True)
    done = 1000
    st.run(g, 1000, 1.0)
    for target in (1000, 1001, 1002, 1003, 1004, 1005, 1040):
        if target > done:
            st.run(g, target - done, 1.0)
            for _ in range(target - done):
                wr += 1.0
                oracle.erosion_iteration3(pts, nb, hr, wr, sr)
            done = target
        out = [torch.empty(g.V, dtype=torch.float64, device="cuda") for _ in range(3)]
        st.store(*out)
        st.ctx.sync_check()
        got = [g.unpermute(t.cpu().numpy()) for t in out]
        print(f"step {target}: h finite {np.isfinite(hr).sum()} inf {np.isinf(hr).sum()} "
              f"nan {np.isnan(hr).sum()}", flush=True)
        for name, x, r in zip("hws", got, (hr, wr, sr)):
            assert _same_bits_or_both_nan(x, r), f"{name} differs at step {target}"
    assert np.isnan(hr).any() and np.isinf(hr).any()


def test_erode_terrain3_snapshots(oracle, tmp_path):
    """erosion.py:186-192: with snapshot, each pass writes the gray image of
    the heights rescaled to 0..255 (then once more by build_image_data,
    util.py:397-398), sampled at each pixel's three nearest vertices and cast
    to uint8, as erosion_snapshot_NNN.png in cfg.SNAP_DIR.  Every image
    equals the oracle's pipeline on the oracle's heights after that pass,
    and the heights equal a run without snapshots."""
    from PIL import Image
    from nixis_amd import cfg
    pts, nb, h = _planet(24)
    W, H = 96, 48
    dists, nbrs = util.KnnGrid(pts).query_pixels_device(W, H, 1.0, 3)
    dn, nn = dists.cpu().numpy(), nbrs.cpu().numpy()
    saved = (cfg.SNAP_DIR, cfg.IMG_QUERY_DATA, cfg.IMG_WIDTH, cfg.IMG_HEIGHT)
    try:
        cfg.SNAP_DIR, cfg.IMG_QUERY_DATA, cfg.IMG_WIDTH, cfg.IMG_HEIGHT = str(tmp_path), (dn, nn), W, H
        hs = h.copy()
        erosion.erode_terrain3(pts, nb, hs, 3, snapshot=True, rain=1.0)
    finally:
        cfg.SNAP_DIR, cfg.IMG_QUERY_DATA, cfg.IMG_WIDTH, cfg.IMG_HEIGHT = saved
    hp = h.copy()
    erosion.erode_terrain3(pts, nb, hp, 3, rain=1.0)
    assert (hs == hp).all()
    href, w, s = h.copy(), np.zeros_like(h), np.zeros_like(h)
    for i in range(3):
        w += 1.0
        oracle.erosion_iteration3(pts, nb, href, w, s)
        layer = oracle.rescale(oracle.rescale(href, 0, 255), 0, 255)
        want = oracle.make_gray_array(W, H, dn, nn, layer).astype(np.uint8)
        got = np.asarray(Image.open(tmp_path / f"erosion_snapshot_{i + 1:03d}.png"))
        assert got.shape == (H, W) and got.dtype == np.uint8
        assert (got == want).all(), f"pass {i + 1}"
    assert (href == hs).all()


def test_erosion_autotune_picks_persists_and_keeps_results(oracle, tmp_path):
    """autotune.tune_erosion times four- and two-step plans (kernel
    variants, then tile chunks), then graph chunks, on a mesh; it stores the
    best for the mesh's size class with the best two-step tuning beside it,
    and lookup_erosion returns them.  Erosion on the tuned graph, and on
    other candidate settings, gives the oracle's heights bit for bit; the
    tuned options hold only for the run and the context's own come back
    after it."""
    from nixis_amd import autotune, batch
    pts, nb, h = _planet(48)
    order = util.sfc_order(pts)
    path = str(tmp_path / "tune.json")
    best, table = autotune.tune_erosion(
        lambda c, k: erosion.ErosionGraph(pts, nb, order=order, chunk=c, steps=k), steps=6,
        reps=1, path=path, variants={4: (0, 1), 2: (0, 1, 5)},
        chunks={4: (256, 512), 2: (256, 512)}, graph_chunks=(8, 32))
    assert len(table) == (2 + 2) + (3 + 2) + 2
    assert {k for _, k, *_ in table} == {2, 4}
    assert best.steps in (2, 4) and best.chunk in (0, 256, 512) and best.graph_chunk in (8, 32)
    assert best.variant in ((0, 1) if best.steps == 4 else (0, 1, 5))
    assert autotune.lookup_erosion(pts.shape[0], path=path) == best
    two = autotune.lookup_erosion(pts.shape[0], path=path, steps=2)
    assert two.steps == 2 and two.variant in (0, 1, 5) and two.chunk in (0, 256, 512)
    href = h.copy()
    oracle.erode_terrain3(pts, nb, href, 21, 1.0)
    ctx = _lib.Context.get()
    tiles = set()
    for kw in (best.graph_kwargs(), two.graph_kwargs(),
               {"chunk": 256, "variant": 5, "graph_chunk": 8, "steps": 2},
               {"chunk": 512, "variant": 1, "graph_chunk": 0, "steps": 4}):
        g = erosion.ErosionGraph(pts, nb, order=order, **kw)
        assert g.steps == kw["steps"]
        tiles.add(g.plan.ntiles)
        st = erosion.ErosionState(g.V)
        st.load(torch.from_numpy(g.permute(h)).cuda())
        st.run(g, 21, 1.0)
        out = torch.empty(g.V, dtype=torch.float64, device="cuda")
        st.store(out)
        st.ctx.sync_check()
        assert (g.unpermute(out.cpu().numpy()) == href).all(), kw
        assert ctx.graph_chunk == 32 and ctx.pair_kernel == 0 and ctx.multi_kernel == 0
    assert len(tiles) >= 2      # the chunks cut different plans
    # MeshCache cuts its plans with a given tuning
    mc = batch.MeshCache()
    g = mc.graph(48, tuning=autotune.ErosionTuning(256, 5, 8, 2))
    assert (g.chunk, g.variant, g.graph_chunk, g.steps) == (256, 5, 8, 2)
    mc.evict(48)
    g = mc.graph(48, tuning=autotune.ErosionTuning(512, 0, 16, 4))
    assert (g.chunk, g.variant, g.graph_chunk, g.steps) == (512, 0, 16, 4)
    # a search of two-step plans alone keeps a faster four-step entry first
    b2, _ = mc.tune_erosion(48, steps=6, reps=1, path=path, variants={2: (0, 5)},
                            chunks={2: (256,)}, graph_chunks=(8,), step_options=(2,))
    assert b2.steps == 2 and mc.graph(48).steps == 2
    assert autotune.lookup_erosion(pts.shape[0], path=path, steps=2) == b2
    if best.steps == 4:
        first = autotune.lookup_erosion(pts.shape[0], path=path)
        assert first.steps == 4 or first == b2
