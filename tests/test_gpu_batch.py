"""Batched planet jobs equal the one-planet-per-call path bit for bit.

Each planet's arithmetic in the batched kernels is the single-planet
kernels' arithmetic, and each planet keeps its own exponent and step
counter, so results must not depend on the batch composition, the tile
size, the blocks per CU, graph capture or chunking of the erosion steps.
"""
import numpy as np
import pytest
import torch

from nixis_amd import autotune, batch

pytestmark = pytest.mark.gpu


def _jobs():
    specs = [(8, 3, 4, 0), (12, 5, 9, 17), (16, 1, 16, 40), (20, 9, 6, 3),
             (30, 2, 11, 0), (24, 77, 13, 25), (10, 4, 5, 40)]
    return [batch.Job(d=d, seed=s, n_octaves=o, erosion_iters=it, rain=1.0 if k % 2 else 0.3 / 320)
            for k, (d, s, o, it) in enumerate(specs)]


def _runs(jobs, meshes):
    return [batch.PlanetRun(j, meshes) for j in jobs]


def _heights(runs):
    torch.cuda.synchronize()
    return [r.elev.cpu().numpy().copy() for r in runs]


@pytest.fixture(scope="module")
def meshes():
    return batch.MeshCache()


@pytest.fixture(scope="module")
def plain(meshes):
    runs = _runs(_jobs(), meshes)
    batch.run_plain(runs)
    runs[0].state  # noqa: B018
    return _heights(runs)


@pytest.mark.parametrize("use_graph,tile,bpc", [(True, 8192, 8), (False, 2048, 2),
                                                (True, 256, 1)])
def test_batch_equals_plain(meshes, plain, use_graph, tile, bpc):
    runs = _runs(_jobs(), meshes)
    br = batch.BatchRunner()
    br.run(runs, batch.Tuning(tile, bpc), use_graph=use_graph)
    br.ctx.sync_check()
    for k, (a, b) in enumerate(zip(_heights(runs), plain)):
        assert (a == b).all(), f"planet {k} differs from the single-planet path"


def test_batch_graph_replays_and_chunks(meshes, plain):
    """Replaying the captured step graph, and running the erosion in chunks
    of 7 steps, give the same heights."""
    runs = _runs(_jobs(), meshes)
    br = batch.BatchRunner()
    for rep in range(3):
        br.run(runs, use_graph=True)
        torch.cuda.synchronize()
        for k, (a, b) in enumerate(zip(_heights(runs), plain)):
            assert (a == b).all(), f"replay {rep} planet {k}"
    br.noise(runs)
    while any(r.state is not None and r.steps_done < r.job.erosion_iters for r in runs):
        br.erode(runs, 7, use_graph=False)
    br.store(runs)
    for k, (a, b) in enumerate(zip(_heights(runs), plain)):
        assert (a == b).all(), f"chunked planet {k}"


@pytest.mark.parametrize("chunk", [5, 6])
def test_batch_one_planet_chunks_under_graphs(meshes, plain, chunk):
    """A planet alone in a pair launch takes the one-planet kernel, which
    reads its step counter on the device: graph-captured chunks of 5 steps
    (two pairs and a single step) or 6 (ending on a pair with no rain after
    it), run twice so the second pass replays the cached graphs at other
    step counts, give the single-planet path's heights."""
    job = _jobs()[2]   # 40 steps
    runs = _runs([job], meshes)
    br = batch.BatchRunner()
    for rep in range(2):
        br.noise(runs)
        while runs[0].steps_done < job.erosion_iters:
            br.erode(runs, chunk, use_graph=True)
        br.store(runs)
        br.ctx.sync_check()
        assert (_heights(runs)[0] == plain[2]).all(), f"pass {rep}"


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("steps", [4, 2])
def test_batch_many_planets_equals_plain(use_graph, steps):
    """More than 16 eroding planets take the many-planet tile kernels, which
    find a tile's planet through a table of buckets over the tiles (tiles
    from chunks of 256 vertices: two tiles per bucket here, so some buckets
    hold the start of a second planet and one-tile planets share buckets);
    with four-step plans the step counts 2..11 give tails of two and three
    steps and single last steps.  Every planet equals its single-planet
    result, with and without graphs."""
    rng = np.random.default_rng(8)
    ds = rng.choice([6, 8, 12, 40, 64, 90, 120], size=24)
    its = rng.integers(2, 12, size=24)
    jobs = [batch.Job(d=int(d), seed=k, n_octaves=4, erosion_iters=int(it), rain=1.0)
            for k, (d, it) in enumerate(zip(ds, its))]
    meshes = batch.MeshCache()
    for d in sorted(set(int(d) for d in ds)):
        g = meshes.graph(d, tuning=autotune.ErosionTuning(256, 0, 32, steps))
        assert g.plan is not None and g.steps == steps
    ref = _runs(jobs, meshes)
    batch.run_plain(ref)
    want = _heights(ref)
    runs = _runs(jobs, meshes)
    br = batch.BatchRunner()
    br.run(runs, use_graph=use_graph)
    br.ctx.sync_check()
    tiles = sum(meshes.graph(j.d).plan.ntiles for j in jobs)
    assert tiles > 1024, tiles
    for k, (a, b) in enumerate(zip(_heights(runs), want)):
        assert (a == b).all(), f"planet {k} (d={jobs[k].d}) differs from the single-planet path"


def test_batch_mixed_plan_steps_equals_plain():
    """One batch holding planets with four-step plans (tails of two and
    three steps, single last steps) and planets with two-step plans: the
    four tile phases and the one-step phase run in one call, and every
    planet equals its single-planet result."""
    meshes = batch.MeshCache()
    four, two = (10, 24, 48), (12, 20, 36)
    for d in four:
        assert meshes.graph(d, tuning=autotune.ErosionTuning(256, 0, 32, 4)).steps == 4
    for d in two:
        assert meshes.graph(d, tuning=autotune.ErosionTuning(256, 0, 32, 2)).steps == 2
    jobs = [batch.Job(d=d, seed=k, n_octaves=5, erosion_iters=it, rain=1.0)
            for k, (d, it) in enumerate(zip(four + two + four, (9, 6, 7, 9, 6, 7, 13, 2, 1)))]
    ref = _runs(jobs, meshes)
    batch.run_plain(ref)
    want = _heights(ref)
    for use_graph in (True, False):
        runs = _runs(jobs, meshes)
        br = batch.BatchRunner()
        br.run(runs, use_graph=use_graph)
        br.ctx.sync_check()
        for k, (a, b) in enumerate(zip(_heights(runs), want)):
            assert (a == b).all(), f"planet {k} (d={jobs[k].d}) differs, graph={use_graph}"


def test_batch_alone_equals_in_mix(meshes, plain):
    """A planet's result is the same alone and inside a larger batch."""
    jobs = _jobs()
    for k in (1, 5):
        runs = _runs([jobs[k]], meshes)
        batch.BatchRunner().run(runs)
        assert (_heights(runs)[0] == plain[k]).all()


def test_batch_matches_oracle(oracle, meshes):
    """Noise within its FP32-gradient tolerance of the oracle; erosion (FP64,
    reference operation order) equal to the oracle's bit for bit when started
    from the batch's own rescaled heights."""
    job = batch.Job(d=14, seed=1234, n_octaves=8, erosion_iters=30, rain=1.0)
    bare = batch.Job(d=14, seed=1234, n_octaves=8, erosion_iters=0, rain=1.0)
    run, run0 = _runs([job, bare], meshes)
    batch.BatchRunner().run([run, run0])
    torch.cuda.synchronize()
    got, h0 = run.heights(), run0.heights()
    from nixis_amd import util
    pts, nb = util.icosphere_with_neighbors(14)
    perm, pgi = oracle.init(1234)
    raw = oracle.sample_octaves(pts, None, perm, pgi, 8)
    h = oracle.rescale(raw, -9000, 9000)
    # each octave's noise is within 1e-5 of FP64 and weighs (n+1)/2 * s_o;
    # the rescale stretches by 18000 / range (min and max move too)
    bound = 3 * 0.5e-5 * sum(0.4 * 0.5 ** o for o in range(8)) * 18000 / (raw.max() - raw.min())
    assert np.abs(h0 - h).max() <= bound
    href = h0.copy()
    oracle.erode_terrain3(pts, nb, href, 30, 1.0)
    assert (got == href).all()


def test_autotune_picks_and_persists(tmp_path, meshes):
    runs = _runs(_jobs()[:4], meshes)
    br = batch.BatchRunner()
    path = str(tmp_path / "tune.json")
    best, table = autotune.tune(br, runs, reps=1, path=path, tiles=(2048, 8192), bpcs=(2, 8))
    assert len(table) == 4 and best.tile in (2048, 8192)
    again = autotune.lookup(runs, path=path)
    assert again == best
