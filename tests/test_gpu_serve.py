"""Served path: every job's result is the same alone and in the mix."""
import numpy as np
import pytest
import torch

from nixis_amd import batch, serve

pytestmark = pytest.mark.gpu


def _trace():
    jobs = serve.served_jobs(14, rate=400.0, seed=11, d_range=(8, 48), iter_range=(20, 60),
                             erosion_frac=0.5)
    return jobs


def test_served_results_equal_alone():
    jobs = _trace()
    srv = serve.PlanetServer(hbm_budget=64 << 20, steps_per_tick=8)
    outs = srv.serve(jobs, keep_results=True)
    torch.cuda.synchronize()
    assert all(not o.rejected for o in outs)
    assert all(o.done >= o.job.arrival for o in outs)
    meshes = batch.MeshCache()
    for o in outs:
        r = batch.PlanetRun(o.job, meshes)
        batch.run_plain([r])
        torch.cuda.synchronize()
        assert (r.elev.cpu().numpy() == o.heights.cpu().numpy()).all(), o.job
    summ = serve.latency_summary(outs)
    assert summ["jobs"] == len(jobs) and summ["p99_ms"] >= summ["p50_ms"] > 0


def test_served_tiny_budget_queues_and_rejects():
    jobs = _trace()
    costs = sorted(serve.job_bytes(j) for j in jobs)
    budget = costs[-2]  # the largest job cannot fit, others queue
    srv = serve.PlanetServer(hbm_budget=budget, steps_per_tick=8)
    outs = srv.serve(jobs)
    rejected = [o for o in outs if o.rejected]
    assert len(rejected) == 1 and serve.job_bytes(rejected[0].job) == costs[-1]
    assert srv.peak_bytes <= budget
    assert all(o.done >= o.admitted >= o.job.arrival - 1e-3 for o in outs if not o.rejected)


@pytest.mark.parametrize("cu_share", [0.5, None])
def test_plain_server_matches_served(cu_share):
    """Served results equal the plain ones, with the noise passes of jobs
    that arrive while others erode on the CU-masked stream (the default) or
    on the whole GPU."""
    jobs = _trace()
    a = serve.PlainServer().serve(jobs, keep_results=True)
    srv = serve.PlanetServer(steps_per_tick=5, noise_cu_share=cu_share)
    if cu_share is None:
        assert srv.noise_masked is None
    else:   # a box that cannot create the masked stream falls back, with the reason kept
        assert (srv.noise_masked is not None) != (srv.noise_mask_error is not None)
    b = srv.serve(jobs, keep_results=True)
    for x, y in zip(a, b):
        assert (x.heights.cpu().numpy() == y.heights.cpu().numpy()).all()


def test_budgeted_ticks_match_plain():
    """Erosion ticks sized by a vertex-step budget (per-planet step caps, a
    budget small enough to split every job over many ticks, and the default
    one), with younger jobs sharing the oldest job's ticks or not and with
    noise passes held to a share of the ticks, give each job the plain
    single-planet result."""
    jobs = _trace()
    a = serve.PlainServer().serve(jobs, keep_results=True)
    for budget, young, noise in ((60_000, 1.0, None), (400_000_000, 1.0, None),
                                 (400_000_000, 0.0, None), (200_000, 0.3, 0.05)):
        b = serve.PlanetServer(tick_vertex_steps=budget, young_share=young,
                               noise_share=noise).serve(jobs, keep_results=True)
        for x, y in zip(a, b):
            assert (x.heights.cpu().numpy() == y.heights.cpu().numpy()).all(), (budget, x.job)
