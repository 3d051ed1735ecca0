"""The measured batched and served workloads, checked on the GPU.

* Batched: the committed 64-planet job list (nixis_amd/jobs/
  batched_64planets.json: d log-uniform in [64, 1024], seeds 0..63, 4..16
  octaves, 0..200 erosion steps) through one BatchRunner run with the
  autotuned schedule and captured erosion graphs.  Every planet equals its
  one-planet-per-call result bit for bit; the three smallest eroding
  planets are checked against the oracle end to end (noise and rescale
  within the FP32-gradient tolerance, erosion bit for bit from the batch's
  own rescaled heights) and the largest planet's noise and rescale against
  the oracle's over the whole planet.
* Served: 24 jobs of the committed 120-job served trace (d log-uniform in
  [32, 2048], including its largest and smallest planets) arriving as a
  Poisson stream while earlier jobs are on the GPU.  Every job equals its
  plain result bit for bit; the smallest eroding job is checked against the
  oracle end to end and the largest noise-only job's heights over the whole
  planet.
"""
from dataclasses import replace

import numpy as np
import pytest
import torch

from nixis_amd import autotune, batch, serve, util

pytestmark = pytest.mark.gpu


def _noise_bound(job, raw):
    """Heights after noise + rescale: each octave's noise is within 1e-5 of
    FP64 and weighs (n + 1) / 2 * s_o; the rescale stretches by
    (upper - lower) / range, and its min and max move by as much."""
    s = sum(job.n_init_strength * job.n_persistence ** o for o in range(job.n_octaves))
    return 3 * 0.5e-5 * s * (job.upper - job.lower) / (raw.max() - raw.min())


def _oracle_heights(oracle, job, pts):
    perm, pgi = oracle.init(job.seed)
    raw = oracle.sample_octaves(pts, None, perm, pgi, job.n_octaves, job.n_init_roughness,
                                job.n_init_strength, job.n_roughness, job.n_persistence,
                                job.world_radius)
    return oracle.rescale(raw, job.lower, job.upper), raw


def _host_order(meshes, d, dev_values):
    out = np.empty_like(dev_values)
    out[meshes.order(d)] = dev_values
    return out


@pytest.fixture(scope="module")
def batched():
    jobs = batch.load_jobs(batch.BATCHED_64_PATH)
    meshes = batch.MeshCache()
    runs = [batch.PlanetRun(j, meshes) for j in jobs]
    br = batch.BatchRunner()
    tuning = autotune.lookup(runs) or batch.Tuning()
    print(f"\n64 planets, {sum(j.V for j in jobs)} vertices, tile {tuning.tile} "
          f"blocks/CU {tuning.blocks_per_cu}", flush=True)
    br.run(runs, tuning, use_graph=True)
    br.run(runs, tuning, use_graph=True)      # a replay of the cached graph
    br.ctx.sync_check()
    got = [r.elev.cpu().numpy() for r in runs]
    return jobs, meshes, runs, got


def test_batched_64_planets_equal_plain(batched):
    jobs, meshes, runs, got = batched
    ref = [batch.PlanetRun(j, meshes) for j in jobs]
    batch.run_plain(ref)
    torch.cuda.synchronize()
    for k, (r, g) in enumerate(zip(ref, got)):
        assert (r.elev.cpu().numpy() == g).all(), f"planet {k} ({jobs[k]}) differs from plain"


def test_batched_64_planets_match_oracle(oracle, batched):
    jobs, meshes, runs, got = batched
    eroding = sorted((k for k, j in enumerate(jobs) if j.erosion_iters > 0),
                     key=lambda k: jobs[k].V)[:3]
    bare_jobs = [replace(jobs[k], erosion_iters=0) for k in eroding]
    bare = [batch.PlanetRun(j, meshes) for j in bare_jobs]
    batch.BatchRunner().run(bare)
    torch.cuda.synchronize()
    for k, b in zip(eroding, bare):
        j = jobs[k]
        pts, nb = util.icosphere_with_neighbors(j.d)
        h_ref, raw = _oracle_heights(oracle, j, pts)
        h0 = b.heights()
        err = np.abs(h0 - h_ref).max()
        assert err <= _noise_bound(j, raw), (j, err)
        href = h0.copy()
        oracle.erode_terrain3(pts, nb, href, j.erosion_iters, j.rain)
        assert (_host_order(meshes, j.d, got[k]) == href).all(), f"planet {k} erosion"
        print(f"planet {k} d={j.d} oct={j.n_octaves} iters={j.erosion_iters}: noise+rescale "
              f"err {err:.2e} m, erosion bit-exact", flush=True)
    big = max(range(len(jobs)), key=lambda k: jobs[k].V)
    j = jobs[big]
    pts = util.create_points(j.d)
    if j.erosion_iters > 0:
        b = batch.PlanetRun(replace(j, erosion_iters=0), meshes)
        batch.BatchRunner().run([b])
        torch.cuda.synchronize()
        h0 = b.heights()
    else:
        h0 = _host_order(meshes, j.d, got[big])
    h_ref, raw = _oracle_heights(oracle, j, pts)
    err = np.abs(h0 - h_ref).max()
    print(f"largest planet {big} d={j.d} V={j.V} oct={j.n_octaves}: noise+rescale err "
          f"{err:.2e} m (bound {_noise_bound(j, raw):.2e})", flush=True)
    assert err <= _noise_bound(j, raw)


def _served_subset():
    full = batch.load_jobs(batch.SERVED_120_PATH)
    by_d = sorted(range(len(full)), key=lambda k: full[k].d)
    pick = set(by_d[:3]) | set(by_d[-3:]) | set(range(0, len(full), 7))
    eroding = [k for k in by_d if full[k].erosion_iters > 0]
    pick |= set(eroding[:2]) | set(eroding[-1:])
    jobs = [full[k] for k in sorted(pick)]
    # Poisson gaps of the trace, rescaled to about 40 jobs per second
    t0 = jobs[0].arrival
    return [replace(j, arrival=(j.arrival - t0) / 40.0) for j in jobs]


@pytest.fixture(scope="module")
def served():
    jobs = _served_subset()
    srv = serve.PlanetServer(hbm_budget=64 << 30)
    print(f"\nserved subset: {len(jobs)} jobs, d {min(j.d for j in jobs)}..{max(j.d for j in jobs)}, "
          f"{sum(j.erosion_iters > 0 for j in jobs)} eroding", flush=True)
    outs = srv.serve(jobs, keep_results=True)
    torch.cuda.synchronize()
    return jobs, srv, outs


def test_served_trace_equals_plain(served):
    jobs, srv, outs = served
    assert all(not o.rejected for o in outs)
    assert max(j.d for j in jobs) >= 1900 and min(j.d for j in jobs) <= 40
    for o in outs:
        r = batch.PlanetRun(o.job, srv.meshes)
        batch.run_plain([r])
        torch.cuda.synchronize()
        assert (r.elev.cpu().numpy() == o.heights.cpu().numpy()).all(), o.job
    summ = serve.latency_summary(outs)
    print(f"served p50 {summ['p50_ms']:.1f} ms p99 {summ['p99_ms']:.1f} ms", flush=True)


def test_served_trace_matches_oracle(oracle, served):
    jobs, srv, outs = served
    small = min((o for o in outs if o.job.erosion_iters > 0), key=lambda o: o.job.V)
    j = small.job
    pts, nb = util.icosphere_with_neighbors(j.d)
    bare = batch.PlanetRun(replace(j, erosion_iters=0), srv.meshes)
    batch.run_plain([bare])
    torch.cuda.synchronize()
    h0 = bare.heights()
    h_ref, raw = _oracle_heights(oracle, j, pts)
    assert np.abs(h0 - h_ref).max() <= _noise_bound(j, raw)
    href = h0.copy()
    oracle.erode_terrain3(pts, nb, href, j.erosion_iters, j.rain)
    got = _host_order(srv.meshes, j.d, small.heights.cpu().numpy())
    assert (got == href).all(), f"served erosion of {j} differs from the oracle"
    big = max((o for o in outs if o.job.erosion_iters == 0), key=lambda o: o.job.V)
    j = big.job
    pts = util.create_points(j.d)
    h_ref, raw = _oracle_heights(oracle, j, pts)
    got = _host_order(srv.meshes, j.d, big.heights.cpu().numpy())
    err = np.abs(got - h_ref).max()
    print(f"served d={j.d} V={j.V} oct={j.n_octaves}: heights err {err:.2e} m "
          f"(bound {_noise_bound(j, raw):.2e})", flush=True)
    assert err <= _noise_bound(j, raw)
