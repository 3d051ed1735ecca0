"""Batched planet jobs: many planets per launch (nx_batch_* in include/nixis.h).

A job is one planet: fractal noise (terrain.sample_octaves with the
reference's defaults unless given), an optional rescale to [lower, upper]
(util.rescale) and an optional number of erosion steps (erode_terrain3).
``BatchRunner`` runs a list of jobs with one persistent noise launch and
erosion launches of four (or two) steps each over the tiles of every planet
still eroding; ``run_plain`` runs
the same jobs one planet per call through the single-planet entry points,
for comparison.  Meshes (vertices and neighbour tables) depend only on the
subdivision count and are shared between jobs through ``MeshCache``.
"""
import ctypes
import json
import os
from dataclasses import asdict, dataclass

import numpy as np

from . import _lib
from . import erosion as _erosion
from . import opensimplex as _osx
from . import terrain as _terrain
from . import util as _util


class NxPlanet(ctypes.Structure):
    """Mirror of nx_planet (include/nixis.h)."""
    _fields_ = [
        ("verts", ctypes.c_void_p), ("elev", ctypes.c_void_p), ("V", ctypes.c_int64),
        ("n_octaves", ctypes.c_int), ("accumulate", ctypes.c_int), ("radius", ctypes.c_double),
        ("rough", ctypes.c_void_p), ("strength", ctypes.c_void_p),
        ("perm", ctypes.c_void_p), ("pgi", ctypes.c_void_p),
        ("lower", ctypes.c_double), ("upper", ctypes.c_double),
        ("erosion_state", ctypes.c_void_p), ("col", ctypes.c_void_p), ("elen", ctypes.c_void_p),
        ("plan", ctypes.c_void_p), ("erosion_iters", ctypes.c_int64), ("rain", ctypes.c_double),
        ("steps_done", ctypes.c_int64), ("step_cap", ctypes.c_int64),
    ]


_SIGS = {
    "nx_batch_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int]),
    "nx_batch_noise": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(NxPlanet), ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_void_p]),
    "nx_batch_erode": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(NxPlanet), ctypes.c_int,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_void_p]),
    "nx_batch_store": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(NxPlanet), ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_void_p]),
}

def _L():
    L = _lib.lib()
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def even_divisions(d):
    """nixis.py:142-146: odd subdivision counts are rounded up to even."""
    d = int(d)
    return d if d % 2 == 0 else d + 1


@dataclass
class Job:
    """One planet job (reference parameter names, defaults from terrain.py:32)."""
    d: int
    seed: int
    n_octaves: int
    erosion_iters: int = 0
    n_init_roughness: float = 1.5
    n_init_strength: float = 0.4
    n_roughness: float = 2.0
    n_persistence: float = 0.5
    world_radius: float = 1.0
    lower: float = -9000.0
    upper: float = 9000.0
    rain: float = _erosion.RAIN_DEFAULT
    arrival: float = 0.0       # seconds, for served traces
    name: str = ""

    @property
    def V(self):
        return 10 * self.d * self.d + 2

    @property
    def vertex_octaves(self):
        return self.V * self.n_octaves

    @property
    def vertex_iters(self):
        return self.V * self.erosion_iters


class MeshCache:
    """Device vertices (and neighbour tables with tile plans when needed) per
    subdivision, numbered in space-filling-curve order (util.sfc_order):
    a planet's fields are per vertex in that order, ``order(d)[k]`` being
    the generator's index of vertex k (``PlanetRun.heights`` maps back).
    With ``on_device`` (default) the icosphere, its neighbour tables, the
    order and the device neighbour layout are generated on the GPU (the same
    arrays as the host generator); the tile plan is cut on the host."""

    def __init__(self, device=None, on_device=True):
        import torch
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.on_device = bool(on_device)
        self._verts = {}
        self._graphs = {}
        self._orders = {}
        self._orders_dev = {}

    def _order_dev(self, d):
        import torch
        od = self._orders_dev.get(d)
        if od is None:
            if self.on_device:
                pts, _ = _util.create_mesh_device(d, self.device, cells=False)
                od = _util.sfc_order_device(pts)
            else:
                od = torch.from_numpy(self.order(d)).to(self.device)
            self._orders_dev[d] = od
        return od

    def order(self, d):
        o = self._orders.get(d)
        if o is None:
            if self.on_device:
                o = self._order_dev(d).cpu().numpy()
            else:
                o = _util.sfc_order(_util.create_points(d))
            self._orders[d] = o
        return o

    def verts(self, d):
        import torch
        v = self._verts.get(d)
        if v is None:
            if self.on_device:
                pts, _ = _util.create_mesh_device(d, self.device, cells=False)
                v = pts.index_select(0, self._order_dev(d)).contiguous()
            else:
                pts = _util.create_points(d)
                v = torch.from_numpy(np.ascontiguousarray(pts[self.order(d)])).to(self.device)
            self._verts[d] = v
        return v

    def graph(self, d, tuning=None):
        """The erosion graph of subdivision d, its tile plan cut with the
        autotuned settings for its size class when the tuning cache has them
        (`tuning` overrides; the results are the same either way)."""
        g = self._graphs.get(d)
        if g is None:
            if tuning is None:
                from . import autotune
                tuning = autotune.lookup_erosion(10 * d * d + 2)
            kw = tuning.graph_kwargs() if tuning is not None else {}
            g = self._build_graph(d, **kw)
            self._graphs[d] = g
        return g

    def _build_graph(self, d, **kw):
        if self.on_device:
            pts, cells = _util.create_mesh_device(d, self.device)
            adj = _util.sort_adjacency_device(_util.build_adjacency_device(cells, pts.shape[0]))
            del cells
            return _erosion.ErosionGraph.from_device(pts, adj, order=self._order_dev(d), **kw)
        pts, nb = _util.icosphere_with_neighbors(d)
        return _erosion.ErosionGraph(pts, nb, order=self.order(d), device=self.device, **kw)

    def tune_erosion(self, d, **kw):
        """Autotune erosion for subdivision d's size class (autotune.tune_erosion)
        and keep the tuned graph; returns (tuning, table)."""
        from . import autotune
        best, table = autotune.tune_erosion(
            lambda c, k: self._build_graph(d, chunk=c, steps=k), **kw)
        self._graphs[d] = self._build_graph(d, **best.graph_kwargs())
        return best, table

    def evict(self, d):
        self._verts.pop(d, None)
        self._graphs.pop(d, None)
        self._orders.pop(d, None)
        self._orders_dev.pop(d, None)

    def nbytes(self):
        n = sum(v.numel() * 8 for v in self._verts.values())
        return n + sum(g.nbytes for g in self._graphs.values())


class PlanetRun:
    """Device buffers and host tables of one job inside a batch."""

    def __init__(self, job, meshes):
        import torch
        self.job = job
        self.verts = meshes.verts(job.d)
        self.V = self.verts.shape[0]
        self.elev = torch.empty(self.V, dtype=torch.float64, device=meshes.device)
        self.perm, self.pgi = _osx.init(job.seed)
        self.rough, self.strength = _terrain.octave_scalars(
            job.n_octaves, job.n_init_roughness, job.n_init_strength, job.n_roughness,
            job.n_persistence, job.world_radius)
        self.meshes = meshes
        self.state = None
        self.graph = None
        if job.erosion_iters > 0:
            self.graph = meshes.graph(job.d)
            self.state = _erosion.ErosionState(self.V, meshes.device)

    @property
    def order(self):
        """order[k] = the generator's index of device vertex k."""
        return self.meshes.order(self.job.d)

    def heights(self):
        """The planet's heights on the host, in the generator's vertex order."""
        out = np.empty(self.V, dtype=np.float64)
        out[self.order] = self.elev.cpu().numpy()
        return out

    def nbytes(self):
        return self.elev.numel() * 8 + (self.state.nbytes if self.state is not None else 0)

    def fill(self, p):
        j = self.job
        p.verts = self.verts.data_ptr()
        p.elev = self.elev.data_ptr()
        p.V = self.V
        p.n_octaves = j.n_octaves
        p.accumulate = 0
        p.radius = float(j.world_radius)
        p.rough = self.rough.ctypes.data
        p.strength = self.strength.ctypes.data
        p.perm = self.perm.ctypes.data
        p.pgi = self.pgi.ctypes.data
        p.lower = float(j.lower)
        p.upper = float(j.upper)
        if self.state is not None:
            p.erosion_state = self.state.buf.data_ptr()
            p.col = self.graph.col.data_ptr()
            p.elen = self.graph.elen.data_ptr()
            p.plan = ctypes.addressof(self.graph.plan) if self.graph.plan is not None else None
            p.erosion_iters = int(j.erosion_iters)
            p.rain = float(j.rain)
        else:
            p.erosion_state = None
            p.col = None
            p.elen = None
            p.plan = None
            p.erosion_iters = 0
            p.rain = 0.0


@dataclass
class Tuning:
    tile: int = 8192
    blocks_per_cu: int = 8


class BatchRunner:
    """Runs batches of PlanetRuns on one stream with one workspace."""

    def __init__(self, device=None, stream=None):
        import torch
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.stream = stream
        self.ctx = _lib.Context.get(self.device)
        self._ws = None

    def _workspace(self, n):
        """Descriptor workspace, allocated on this runner's stream so the
        caching allocator only reuses it in that stream's order (the kernels
        of earlier calls on the stream may still be reading it)."""
        import torch
        need = int(_L().nx_batch_workspace_bytes(n))
        if self._ws is None or self._ws.numel() < need:
            with torch.cuda.stream(self.stream or torch.cuda.current_stream(self.device)):
                self._ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        return self._ws

    @staticmethod
    def _structs(runs):
        arr = (NxPlanet * len(runs))()
        for i, r in enumerate(runs):
            r.fill(arr[i])
        return arr

    def noise(self, runs, tuning=None):
        """Noise (+ rescale + erosion state load) for every run."""
        t = tuning or Tuning()
        arr = self._structs(runs)
        ws = self._workspace(len(runs))
        _lib.check(_L().nx_batch_noise(self.ctx.handle, arr, len(runs), ws.data_ptr(),
                                       int(t.tile), int(t.blocks_per_cu),
                                       _lib.stream_ptr(self.stream)))
        for r in runs:
            r.steps_done = 0
        return arr

    def erode(self, runs, max_steps, use_graph=True, caps=None):
        """Up to max_steps erosion steps for every run with steps left
        (caps: optional per-run limits for this call)."""
        arr = self._structs(runs)
        for i, r in enumerate(runs):
            arr[i].steps_done = getattr(r, "steps_done", 0)
            arr[i].step_cap = int(caps[i]) if caps is not None else 0
        ws = self._workspace(len(runs))
        _lib.check(_L().nx_batch_erode(self.ctx.handle, arr, len(runs), int(max_steps),
                                       ws.data_ptr(), 1 if use_graph else 0,
                                       _lib.stream_ptr(self.stream)))
        for i, r in enumerate(runs):
            r.steps_done = arr[i].steps_done

    def store(self, runs):
        """Eroded heights into each run's elev."""
        runs = [r for r in runs if r.state is not None]
        if not runs:
            return
        arr = self._structs(runs)
        ws = self._workspace(len(runs))
        _lib.check(_L().nx_batch_store(self.ctx.handle, arr, len(runs), ws.data_ptr(),
                                       _lib.stream_ptr(self.stream)))

    def run(self, runs, tuning=None, use_graph=True):
        """The whole batch: noise, rescale, all erosion steps, store."""
        self.noise(runs, tuning)
        max_iters = max((r.job.erosion_iters for r in runs), default=0)
        if max_iters > 0:
            self.erode(runs, max_iters, use_graph)
            self.store(runs)


def run_plain(runs, stream=None):
    """The same jobs one planet per call through the single-planet API."""
    for r in runs:
        j = r.job
        _terrain.sample_octaves_device(r.verts, None, r.perm, r.pgi, r.rough, r.strength,
                                       j.world_radius, stream=stream, out=r.elev)
        if j.lower < j.upper:
            _util.rescale_device(r.elev, j.lower, j.upper, out=r.elev, stream=stream)
        if r.state is not None:
            r.state.load(r.elev, stream=stream)
            r.state.run(r.graph, j.erosion_iters, j.rain, stream=stream)
            r.state.store(r.elev, stream=stream)


JOBS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jobs")
BATCHED_64_PATH = os.path.join(JOBS_DIR, "batched_64planets.json")
SERVED_120_PATH = os.path.join(JOBS_DIR, "served_120jobs_rate1.json")


def save_jobs(jobs, path):
    """Write a job list as JSON (one object per job, every Job field)."""
    with open(path, "w") as f:
        f.write("[\n" + ",\n".join(json.dumps(asdict(j)) for j in jobs) + "\n]\n")


def load_jobs(path):
    """Read a job list written by save_jobs."""
    with open(path) as f:
        return [Job(**e) for e in json.load(f)]


def batched_jobs(n=64, seed=20261202, d_range=(64, 1024), oct_range=(4, 16),
                 iter_range=(0, 200)):
    """Batched-form job list: d log-uniform in d_range (made even), seeds
    0..n-1, octaves uniform in oct_range, erosion steps uniform in
    iter_range, from numpy.random.default_rng(seed)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(d_range[0]), np.log(d_range[1])
    jobs = []
    for i in range(n):
        d = even_divisions(int(round(float(np.exp(rng.uniform(lo, hi))))))
        o = int(rng.integers(oct_range[0], oct_range[1] + 1))
        it = int(rng.integers(iter_range[0], iter_range[1] + 1))
        jobs.append(Job(d=d, seed=i, n_octaves=o, erosion_iters=it, name=f"b{i:02d}"))
    return jobs
