"""Served planet generation: continuous batching under an HBM budget.

Jobs (nixis_amd.batch.Job) arrive over time.  ``Admission`` decides, in
arrival order, which jobs fit the HBM budget for per-job device fields: the
FP64 heights and, for a job that erodes, its FP64 erosion state.
Meshes (vertices and neighbour tables) depend only on the subdivision count;
they are shared assets provisioned before the trace starts.  A job whose
fields exceed the whole budget is rejected; others wait, first come first
served, until enough of the budget is released by finished jobs.

``PlanetServer.serve`` runs a trace in real time.  Each tick launches one
batched noise pass (on the noise stream) for the jobs admitted since the
last tick and one batched chunk of erosion steps (on the erosion stream),
so jobs that arrive while others are on the GPU join the next tick
(continuous batching) and noise-only jobs do not wait behind erosion; while
jobs erode, noise passes may use only part of the CUs (noise_cu_share).  An
erosion tick hands out a budget of vertex-steps to the eroding jobs oldest
first, each job taking as many of its remaining steps as the budget left
covers, in whole launches of its tile plan (four steps, or two; at least
one launch): a small planet finishes all its steps in one tick
while a large one advances a few steps per tick in arrival order.  By
default the oldest eroding job has each tick to itself (young_share) unless
its planet is too small to fill the GPU (fill_vertices), and noise passes may
be held to a share of the tick's time (noise_share).  At most a few ticks are
in flight.  Completion times come from GPU events relative to an event
recorded at the start of the trace.
Results stay on the device.  ``PlainServer`` runs the same trace one planet
per call, in arrival order, for comparison.
"""
import os
import sys
import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import batch as _batch


def _launch_steps(run):
    """Steps per tile launch of a run's erosion graph (4 or 2; 2 without a
    plan or a graph)."""
    g = getattr(run, "graph", None)
    if g is None or getattr(g, "plan", None) is None:
        return 2
    return int(g.steps)


def job_bytes(job):
    """Device bytes of a job's own fields."""
    from ._lib import lib
    b = 8 * job.V
    if job.erosion_iters > 0:
        b += int(lib().nx_erosion_state_bytes(job.V))
    return b


class Admission:
    """First-come-first-served admission against an HBM budget."""

    def __init__(self, budget, job_cost=job_bytes):
        self.budget = int(budget)
        self.job_cost = job_cost
        self.used = 0
        self.peak = 0

    def fits_ever(self, job):
        return self.job_cost(job) <= self.budget

    def try_admit(self, job):
        need = self.job_cost(job)
        if self.used + need > self.budget:
            return False
        self.used += need
        self.peak = max(self.peak, self.used)
        return True

    def release(self, job):
        self.used -= self.job_cost(job)
        assert self.used >= 0


@dataclass
class Outcome:
    job: _batch.Job
    admitted: float = float("nan")
    done: float = float("nan")
    rejected: bool = False
    heights: object = None   # device tensor (served) or None

    @property
    def latency(self):
        return self.done - self.job.arrival


def latency_summary(outcomes, work_start=0.0):
    lat = np.array([o.latency for o in outcomes if not o.rejected])
    done = max(o.done for o in outcomes if not o.rejected)
    vo = sum(o.job.vertex_octaves for o in outcomes if not o.rejected)
    vi = sum(o.job.vertex_iters for o in outcomes if not o.rejected)
    span = done - work_start
    return {
        "jobs": int(lat.size),
        "rejected": int(sum(o.rejected for o in outcomes)),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
        "makespan_s": float(span),
        "jobs_per_s": float(lat.size / span),
        "Gvertex_octaves_per_s": float(vo / span / 1e9),
        "Gvertex_iters_per_s": float(vi / span / 1e9),
    }


class _Clock:
    """Host clock and GPU event timeline sharing one origin."""

    def __init__(self, stream):
        import torch
        self.stream = stream
        self.t0_event = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        self.t0_event.record(stream)
        self.t0_event.synchronize()
        self.t0 = time.perf_counter()

    def now(self):
        return time.perf_counter() - self.t0

    def event(self, stream=None):
        import torch
        e = torch.cuda.Event(enable_timing=True)
        e.record(stream or self.stream)
        return e

    def gpu_time(self, e):
        return self.t0_event.elapsed_time(e) * 1e-3


def _prepare_meshes(meshes, jobs):
    """Per-subdivision assets (vertices, neighbour tables) built up front."""
    for j in jobs:
        meshes.verts(j.d)
        meshes.order(j.d)
        if j.erosion_iters > 0:
            meshes.graph(j.d)


# Rates of the device kernels used to size a tick's noise admissions (MI355X:
# fused noise about 2.3e11 vertex-octaves/s, four-step erosion about 1.2e11
# vertex-steps/s); only their ratio matters.
NOISE_RATE = 2.3e11
EROSION_RATE = 1.2e11


class PlanetServer:
    def __init__(self, hbm_budget=64 << 30, steps_per_tick=None, tuning=None, device=None,
                 tick_vertex_steps=400_000_000, noise_share=None, young_share=0.0,
                 max_queued=3, noise_cu_share=0.5, fill_vertices=None):
        import torch
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.stream = torch.cuda.current_stream(self.device)
        self.erode_stream = _lib.side_stream(self.device)
        # While jobs erode, noise passes run on a stream whose kernels may use
        # only noise_cu_share of the CUs (a CU mask; None or 1 for all): a
        # kernel trace of the 120-job trace showed the erosion launches that
        # overlap a noise pass running up to twice as long when the noise
        # blocks may take every CU (the 64 KB erosion blocks then find no
        # room), and the large jobs at the tail finishing about 5 ms later.
        self.noise_masked = None
        self.noise_mask_error = None
        if noise_cu_share is not None and 0.0 < noise_cu_share < 1.0:
            ncu = torch.cuda.get_device_properties(self.device).multi_processor_count
            try:
                self.noise_masked = _lib.masked_stream(max(1, int(ncu * noise_cu_share)),
                                                       self.device)
            except _lib.NixisError as e:
                # a scheduling choice, not a result: without the mask the
                # noise passes use every CU (recorded for the caller)
                self.noise_mask_error = str(e)
        # at most this many ticks (noise passes and erosion chunks) queued on
        # the GPU before the host waits for the oldest
        self.max_queued = max(1, int(max_queued))
        # a tick whose planets have fewer vertices than this leaves the GPU
        # short of work, so younger jobs join it (about four 1024-vertex tiles
        # per CU; measured on the 120-job trace with every job at t=0: 0 gave
        # 0.565-0.568 s, two 768-vertex tiles per CU 0.562, 1 M vertices
        # 0.561, 4 M 0.562, against 0.564 one planet per call)
        if fill_vertices is None:
            ncu = torch.cuda.get_device_properties(self.device).multi_processor_count
            fill_vertices = 4 * 1024 * ncu
        self.fill_vertices = int(fill_vertices)
        self.meshes = _batch.MeshCache(self.device)
        self.noise_runner = _batch.BatchRunner(self.device, self.stream)
        self.masked_runner = (_batch.BatchRunner(self.device, self.noise_masked)
                              if self.noise_masked is not None else None)
        self.erode_runner = _batch.BatchRunner(self.device, self.erode_stream)
        self.budget = int(hbm_budget)
        # steps_per_tick: a fixed per-job cap per tick instead of the budget
        self.steps_per_tick = int(steps_per_tick) if steps_per_tick else 0
        self.tick_vertex_steps = int(tick_vertex_steps)
        self.tuning = tuning
        # noise_share: while jobs erode, a tick starts new jobs (first fit in
        # arrival order, at least one) whose noise takes at most this share
        # of the tick's erosion time; None starts every job that fits.
        self.noise_share = None if noise_share is None else float(noise_share)
        # young_share: of a tick's vertex-steps, at most this share goes to
        # eroding jobs other than the oldest.  0 (the default) erodes the
        # jobs in arrival order, the oldest alone in its tick unless it is
        # too small to fill the GPU (fill_vertices): measured on the 120-job
        # trace, giving younger jobs the oldest one's leftover budget delayed
        # the largest jobs (p99 163.6 ms at 1.0, 156.4 ms at 0)
        self.young_share = float(young_share)
        self.trace = bool(os.environ.get("NX_SERVE_TRACE"))

    def _tick(self, eroding):
        """(jobs, step caps) of the next erosion tick: oldest first, each
        taking the steps the remaining budget covers, in whole launches of
        its tile plan's steps (at least one launch).  Younger
        jobs share the tick's budget up to young_share of it, and beyond that
        while the tick's planets together have fewer than fill_vertices
        vertices (too few tiles to fill the GPU), for at most the oldest
        job's number of steps."""
        tick, caps, left = [], [], self.tick_vertex_steps
        young_left = int(self.young_share * self.tick_vertex_steps)
        width = 0
        for r in eroding:
            allow = left
            if tick:
                allow = min(left, young_left)
                if width < self.fill_vertices:
                    allow = max(allow, min(left, caps[0] * r.V))
                if allow < 2 * r.V:
                    break
            remaining = r.job.erosion_iters - r.steps_done
            n = remaining if self.steps_per_tick == 0 else min(remaining, self.steps_per_tick)
            n = min(n, max(2, allow // r.V))
            unit = _launch_steps(r)
            if n < remaining and n % unit:
                # whole launches; a shorter one only at the end, and never
                # fewer than one launch, so every tick advances
                n = max(n - n % unit, min(unit, remaining))
            if tick:
                young_left -= n * r.V
            tick.append(r)
            caps.append(n)
            left -= n * r.V
            width += r.V
        return tick, caps

    def _admit(self, waiting, adm, eroding):
        """Jobs to start now, removed from `waiting`.  Admission against the
        HBM budget is first come first served (the first job that does not
        fit stops it).  With noise_share set and jobs eroding, the noise of
        the jobs started is kept to noise_share times the estimated time of
        the next erosion tick: jobs over that budget wait for a later tick
        while smaller ones behind them may start (first fit), and the first
        job always starts, so every tick admits work."""
        if self.noise_share is None or not eroding:
            out = []
            while waiting and adm.try_admit(waiting[0]):
                out.append(waiting.popleft())
            return out
        tick, caps = self._tick(eroding)
        budget = self.noise_share * sum(c * r.V for r, c in zip(tick, caps)) / EROSION_RATE
        out, later, spent = [], deque(), 0.0
        while waiting:
            j = waiting[0]
            cost = j.vertex_octaves / NOISE_RATE
            if out and spent + cost > budget:
                later.append(waiting.popleft())
                continue
            if not adm.try_admit(j):
                break
            out.append(waiting.popleft())
            spent += cost
        waiting.extendleft(reversed(later))
        return out

    def serve(self, jobs, keep_results=False, prepare_meshes=True):
        jobs = sorted(jobs, key=lambda j: j.arrival)
        if prepare_meshes:
            _prepare_meshes(self.meshes, jobs)
        adm = Admission(self.budget)
        outcomes = {id(j): Outcome(j) for j in jobs}
        pending = deque(jobs)
        waiting = deque()
        eroding = []
        inflight = deque()   # (event, [runs completed by it])
        clock = _Clock(self.stream)
        while pending or waiting or eroding or inflight:
            now = clock.now()
            while pending and pending[0].arrival <= now:
                j = pending.popleft()
                if adm.fits_ever(j):
                    waiting.append(j)
                else:
                    outcomes[id(j)].rejected = True
            admitted = []
            for j in self._admit(waiting, adm, eroding):
                outcomes[id(j)].admitted = now
                admitted.append(_batch.PlanetRun(j, self.meshes))
            if admitted:
                if eroding and self.masked_runner is not None:
                    # the fields were allocated in the current stream's order
                    ns = self.noise_masked
                    ns.wait_stream(self.stream)
                    for r in admitted:
                        r.elev.record_stream(ns)
                        if r.state is not None:
                            r.state.buf.record_stream(ns)
                    self.masked_runner.noise(admitted, self.tuning)
                else:
                    ns = self.stream
                    self.noise_runner.noise(admitted, self.tuning)
                ev = clock.event(ns)
                done_now = [r for r in admitted if r.state is None]
                inflight.append((ev, done_now))
                new_erode = [r for r in admitted if r.state is not None]
                if new_erode:
                    self.erode_stream.wait_event(ev)
                    for r in new_erode:
                        # allocated in the current stream's order, used on the
                        # erosion stream: keep the allocator from reusing them early
                        r.elev.record_stream(self.erode_stream)
                        r.state.buf.record_stream(self.erode_stream)
                    eroding += new_erode
            if eroding:
                tick, caps = self._tick(eroding)
                self.erode_runner.erode(tick, max(caps), use_graph=False, caps=caps)
                finished = [r for r in tick if r.steps_done >= r.job.erosion_iters]
                if finished:
                    self.erode_runner.store(finished)
                    eroding = [r for r in eroding if r.steps_done < r.job.erosion_iters]
                inflight.append((clock.event(self.erode_stream), finished))
            if self.trace:
                print(f"[serve] t={now:.4f} pending={len(pending)} waiting={len(waiting)} "
                      f"admitted={len(admitted)} eroding={len(eroding)} "
                      f"inflight={len(inflight)} used={adm.used}", file=sys.stderr, flush=True)
            # Keep at most max_queued ticks queued; retire what has finished.
            while inflight and (len(inflight) > self.max_queued or inflight[0][0].query()):
                ev, done_runs = inflight.popleft()
                ev.synchronize()
                t = clock.gpu_time(ev)
                for r in done_runs:
                    o = outcomes[id(r.job)]
                    o.done = t
                    if keep_results:
                        o.heights = r.elev
                    adm.release(r.job)
            if not (admitted or eroding) and pending and not waiting and not inflight:
                time.sleep(max(0.0, min(0.002, pending[0].arrival - clock.now())))
            elif not (admitted or eroding) and inflight:
                time.sleep(0.0002)
        self.peak_bytes = adm.peak
        return [outcomes[id(j)] for j in jobs]


class PlainServer:
    """One planet per call, in arrival order, through the single-planet API."""

    def __init__(self, device=None):
        import torch
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.stream = torch.cuda.current_stream(self.device)
        self.meshes = _batch.MeshCache(self.device)

    def serve(self, jobs, keep_results=False, prepare_meshes=True):
        jobs = sorted(jobs, key=lambda j: j.arrival)
        if prepare_meshes:
            _prepare_meshes(self.meshes, jobs)
        outcomes = []
        marks = []
        clock = _Clock(self.stream)
        for j in jobs:
            wait = j.arrival - clock.now()
            if wait > 0:
                time.sleep(wait)
            o = Outcome(j, admitted=max(j.arrival, clock.now()))
            r = _batch.PlanetRun(j, self.meshes)
            _batch.run_plain([r], stream=self.stream)
            marks.append((clock.event(), o, r))
            outcomes.append(o)
            # retire finished jobs so their buffers can be freed
            while marks and marks[0][0].query():
                ev, oo, rr = marks.pop(0)
                oo.done = clock.gpu_time(ev)
                if keep_results:
                    oo.heights = rr.elev
        for ev, oo, rr in marks:
            ev.synchronize()
            oo.done = clock.gpu_time(ev)
            if keep_results:
                oo.heights = rr.elev
        return outcomes


def served_jobs(n, rate, seed=20261203, d_range=(32, 2048), oct_range=(4, 16),
                erosion_frac=0.3, iter_range=(100, 300)):
    """Served-form trace: Poisson arrivals at `rate` jobs/s, d log-uniform in
    d_range (made even), random seeds, octaves uniform in oct_range, and a
    fraction `erosion_frac` of jobs with erosion steps uniform in iter_range.
    Job attributes and arrival gaps come from separate generators seeded by
    `seed`, so changing `rate` only rescales the arrival times."""
    rng = np.random.default_rng(seed)
    gaps = np.random.default_rng(seed + 1).exponential(1.0, size=n) / rate
    lo, hi = np.log(d_range[0]), np.log(d_range[1])
    t = 0.0
    jobs = []
    for i in range(n):
        t += float(gaps[i])
        d = _batch.even_divisions(int(round(float(np.exp(rng.uniform(lo, hi))))))
        o = int(rng.integers(oct_range[0], oct_range[1] + 1))
        erode = bool(rng.uniform() < erosion_frac)
        it = int(rng.integers(iter_range[0], iter_range[1] + 1)) if erode else 0
        seed_i = int(rng.integers(0, 2**31 - 1))
        jobs.append(_batch.Job(d=d, seed=seed_i, n_octaves=o, erosion_iters=it, arrival=t,
                               name=f"s{i:03d}"))
    return jobs


def load_trace(path, rate):
    """A served trace saved at one job per second (batch.save_jobs of
    served_jobs(n, rate=1.0)) with its arrival times rescaled to `rate`."""
    from dataclasses import replace
    return [replace(j, arrival=j.arrival / rate) for j in _batch.load_jobs(path)]
