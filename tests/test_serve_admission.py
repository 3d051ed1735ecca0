"""Admission control of the served path (host logic, CPU only)."""
import pytest

from nixis_amd.batch import Job
from nixis_amd.serve import Admission


def cost(job):
    return job.V  # one "byte" per vertex keeps the arithmetic readable


def test_fifo_budget_never_exceeded():
    adm = Admission(budget=10_000, job_cost=cost)
    jobs = [Job(d=d, seed=0, n_octaves=4) for d in (20, 10, 25, 8, 30)]
    # V: 4002, 1002, 6252, 642, 9002
    admitted = [j for j in jobs if adm.try_admit(j)]
    assert [j.d for j in admitted] == [20, 10, 8]  # 25 does not fit after 20+10
    assert adm.used <= adm.budget
    # try_admit never over-commits, and releases free the budget
    adm.release(jobs[0])
    adm.release(jobs[1])
    assert adm.try_admit(jobs[2])
    assert adm.used == 642 + 6252 and adm.peak <= adm.budget


def test_job_larger_than_budget_is_rejected_up_front():
    adm = Admission(budget=5_000, job_cost=cost)
    big = Job(d=30, seed=1, n_octaves=4)
    assert not adm.fits_ever(big)
    assert not adm.try_admit(big)
    assert adm.used == 0


def test_waiting_job_admitted_once_budget_frees():
    adm = Admission(budget=7_000, job_cost=cost)
    a = Job(d=20, seed=0, n_octaves=4)   # 4002
    b = Job(d=18, seed=0, n_octaves=4)   # 3242
    assert adm.try_admit(a)
    assert not adm.try_admit(b)          # waits
    adm.release(a)
    assert adm.try_admit(b)


def test_release_below_zero_is_a_bug():
    adm = Admission(budget=100, job_cost=cost)
    with pytest.raises(AssertionError):
        adm.release(Job(d=2, seed=0, n_octaves=1))


def test_served_trace_shape():
    from nixis_amd.serve import served_jobs
    jobs = served_jobs(400, rate=50.0, seed=7)
    arr = [j.arrival for j in jobs]
    assert arr == sorted(arr)
    assert all(32 <= j.d <= 2048 and j.d % 2 == 0 for j in jobs)
    assert all(4 <= j.n_octaves <= 16 for j in jobs)
    frac = sum(j.erosion_iters > 0 for j in jobs) / len(jobs)
    assert 0.2 < frac < 0.4
    assert all(100 <= j.erosion_iters <= 300 for j in jobs if j.erosion_iters)
    # mean inter-arrival ~ 1/rate
    gaps = [b - a for a, b in zip([0.0] + arr[:-1], arr)]
    assert abs(sum(gaps) / len(gaps) - 1 / 50.0) < 0.005


class _Plan:
    pass


class _Graph:
    def __init__(self, steps):
        self.steps = steps
        self.plan = _Plan()


class _Run:
    def __init__(self, V, iters, steps=None):
        self.V = V
        self.steps_done = 0
        self.job = Job(d=2, seed=0, n_octaves=1, erosion_iters=iters)
        if steps is not None:
            self.graph = _Graph(steps)


@pytest.mark.parametrize("unit", [2, 4])
@pytest.mark.parametrize("budget, per_tick", [(9_000, 0), (4_000, 0), (50_000, 5), (50_000, 1),
                                              (50_000, 3)])
def test_erosion_ticks_hand_out_whole_launches(budget, per_tick, unit):
    """Every tick gives a job a whole number of its plan's launches (two
    or four steps) except the tick that finishes it, and the steps add up to
    the job's.  A per-job cap below one launch still advances every job by
    a launch per tick (rounding down never reaches zero steps)."""
    from nixis_amd.serve import PlanetServer
    srv = PlanetServer.__new__(PlanetServer)
    srv.tick_vertex_steps, srv.steps_per_tick, srv.young_share = budget, per_tick, 1.0
    srv.fill_vertices = 0
    runs = [_Run(1000, 31, unit), _Run(400, 7, unit), _Run(3000, 10, unit)]
    eroding, ticks = list(runs), 0
    while eroding:
        tick, caps = srv._tick(eroding)
        assert tick and all(n >= 1 for n in caps)
        for r, n in zip(tick, caps):
            left = r.job.erosion_iters - r.steps_done
            assert n <= left and (n % unit == 0 or n == left)
            r.steps_done += n
        eroding = [r for r in eroding if r.steps_done < r.job.erosion_iters]
        ticks += 1
        assert ticks < 100
    assert [r.steps_done for r in runs] == [31, 7, 10]


def test_noise_share_admits_first_fit_within_the_tick_budget():
    """With jobs eroding and a noise share set, a tick starts the waiting
    jobs whose noise fits noise_share times the tick's estimated erosion
    time, first fit in arrival order, and always at least the first; the
    skipped jobs stay waiting in their order.  Without eroding jobs, or
    without a share, every job that fits the HBM budget starts."""
    from collections import deque
    from nixis_amd import serve
    from nixis_amd.serve import PlanetServer
    srv = PlanetServer.__new__(PlanetServer)
    srv.tick_vertex_steps, srv.steps_per_tick, srv.young_share = 10**12, 0, 1.0
    srv.fill_vertices = 0
    eroding = [_Run(1_000_000, 40)]          # tick: 40 steps x 1e6 vertices
    tick_s = 40 * 1_000_000 / serve.EROSION_RATE

    def job(vo_s, name):
        # a job whose noise takes vo_s seconds at NOISE_RATE
        d = 2
        octs = max(1, int(round(vo_s * serve.NOISE_RATE / (10 * d * d + 2))))
        return Job(d=d, seed=0, n_octaves=octs, name=name)

    share = 0.5
    a, b, c, dd = (job(0.6 * share * tick_s, "a"), job(0.6 * share * tick_s, "b"),
                   job(0.3 * share * tick_s, "c"), job(0.05 * share * tick_s, "d"))
    for noise_share, erode, want, left in ((share, eroding, "acd", "b"), (None, eroding, "abcd", ""),
                                           (share, [], "abcd", "")):
        srv.noise_share = noise_share
        waiting = deque([a, b, c, dd])
        adm = Admission(budget=10**12, job_cost=cost)
        got = srv._admit(waiting, adm, erode)
        assert "".join(j.name for j in got) == want
        assert "".join(j.name for j in waiting) == left
    # the first job always starts, even when its noise alone exceeds the budget
    srv.noise_share = share
    waiting = deque([job(10 * tick_s, "x"), a])
    got = srv._admit(waiting, Admission(budget=10**12, job_cost=cost), eroding)
    assert [j.name for j in got] == ["x"] and [j.name for j in waiting] == ["a"]


def test_young_share_keeps_the_oldest_job_alone():
    """young_share 0 gives the oldest eroding job the whole tick; a share
    lets younger jobs use at most that share of the tick's budget."""
    from nixis_amd.serve import PlanetServer
    srv = PlanetServer.__new__(PlanetServer)
    srv.tick_vertex_steps, srv.steps_per_tick, srv.fill_vertices = 100_000, 0, 0
    runs = [_Run(1000, 50), _Run(100, 40), _Run(100, 40)]   # the oldest leaves 50,000
    srv.young_share = 0.0
    tick, caps = srv._tick(runs)
    assert tick == runs[:1] and caps == [50]
    srv.young_share = 0.05            # at most 5,000 vertex-steps for the younger jobs
    tick, caps = srv._tick(runs)
    assert tick == runs and caps == [50, 40, 10]
    srv.young_share = 1.0
    tick, caps = srv._tick(runs)
    assert tick == runs and caps == [50, 40, 40]


def test_ticks_of_small_jobs_are_filled():
    """With young_share 0, younger jobs still join a tick while its planets
    together have fewer than fill_vertices vertices, each for at most the
    oldest job's number of steps; an oldest job that fills the GPU keeps the
    tick to itself."""
    from nixis_amd.serve import PlanetServer
    srv = PlanetServer.__new__(PlanetServer)
    srv.tick_vertex_steps, srv.steps_per_tick, srv.young_share = 100_000, 0, 0.0
    runs = [_Run(1000, 50), _Run(100, 40), _Run(300, 60), _Run(2000, 10)]
    srv.fill_vertices = 1500
    tick, caps = srv._tick(runs)
    assert tick == runs and caps == [50, 40, 50, 10]
    srv.fill_vertices = 1050
    tick, caps = srv._tick(runs)
    assert tick == runs[:2] and caps == [50, 40]
    srv.fill_vertices = 500
    tick, caps = srv._tick(runs)
    assert tick == runs[:1] and caps == [50]
