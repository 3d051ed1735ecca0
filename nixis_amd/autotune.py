"""Autotuner for the noise and erosion launches.

Noise (batched form): tile size x blocks per CU of the persistent launch,
keyed by the batch's shape class (log2 of the total vertex count and the
mean octave count).

Erosion: the tile plan's steps per launch (four or two), its starting
chunk (tile size before halving), the kernel variant and the steps per
captured hipGraph, keyed by the mesh's size class (log2 of its vertex
count).  For each step count a coordinate search times the variants on the
default plan, then the chunks with the best variant; the graph chunks are
timed for the best of all.  None of the choices changes the results.

Candidates are timed with HIP events on the launch stream; the best one is
stored in a JSON file, so later work of the same class reuses it without
timing.
"""
import json
import math
import os
import threading
from dataclasses import dataclass

from .batch import Tuning

CACHE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_cache.json")
TILES = (2048, 4096, 8192, 16384, 32768)
BLOCKS_PER_CU = (2, 4, 8)
_lock = threading.Lock()


def shape_key(runs):
    V = sum(r.V for r in runs)
    mean_oct = sum(r.job.n_octaves * r.V for r in runs) / max(V, 1)
    return f"noise_batch:V2^{int(math.log2(max(V, 1)))}:oct{int(round(mean_oct))}"


def load_cache(path=CACHE_PATH):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache, path=CACHE_PATH):
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def lookup(runs, path=CACHE_PATH):
    """Cached tuning for this batch's class, or None."""
    e = load_cache(path).get(shape_key(runs))
    return Tuning(e["tile"], e["blocks_per_cu"]) if e else None


def tune(runner, runs, reps=3, path=CACHE_PATH, tiles=TILES, bpcs=BLOCKS_PER_CU, persist=True):
    """Time every (tile, blocks_per_cu) on these runs; return the best Tuning
    and the table of timings (ms per batched noise launch)."""
    import torch
    stream = runner.stream or torch.cuda.current_stream()
    results = []
    for tile in tiles:
        for bpc in bpcs:
            t = Tuning(tile, bpc)
            runner.noise(runs, t)  # warm
            start = torch.cuda.Event(enable_timing=True)
            end = torch.cuda.Event(enable_timing=True)
            start.record(stream)
            for _ in range(reps):
                runner.noise(runs, t)
            end.record(stream)
            end.synchronize()
            results.append((start.elapsed_time(end) / reps, tile, bpc))
    results.sort()
    best = Tuning(results[0][1], results[0][2])
    if persist:
        with _lock:
            cache = load_cache(path)
            cache[shape_key(runs)] = {"tile": best.tile, "blocks_per_cu": best.blocks_per_cu,
                                      "ms": round(results[0][0], 4)}
            save_cache(cache, path)
    return best, results


# ---- erosion ---------------------------------------------------------------

PAIR_VARIANTS = (0, 1, 2, 4, 5, 6)     # nx_ctx option NX_OPT_PAIR_KERNEL (3 repeats 0)
MULTI_VARIANTS = (0,)                  # NX_OPT_MULTI_KERNEL (1..4 are slower everywhere measured)
VARIANTS = {4: MULTI_VARIANTS, 2: PAIR_VARIANTS}
CHUNKS = {4: (896, 1024, 1088, 1216, 1280), 2: (512, 640, 704, 768, 832, 896)}
STEPS = (4, 2)
GRAPH_CHUNKS = (8, 16, 32, 64)


@dataclass
class ErosionTuning:
    chunk: int = 0          # tile plan starting chunk (0: the library default)
    variant: int = 0        # kernel variant for the plan's steps per launch
    graph_chunk: int = 32   # erosion steps per captured graph
    steps: int = 4          # steps per launch of the tile plan (4 or 2)

    def graph_kwargs(self):
        return {"chunk": self.chunk, "variant": self.variant, "graph_chunk": self.graph_chunk,
                "steps": self.steps}


def erosion_key(V):
    return f"erosion:V2^{int(math.log2(max(int(V), 1)))}"


def _tuning(e):
    return ErosionTuning(e["chunk"], e["variant"], e["graph_chunk"], e.get("steps", 2))


def lookup_erosion(V, path=CACHE_PATH, steps=None):
    """Cached erosion tuning for meshes of V vertices' size class, or None;
    with steps=2 (or 4), the best tuning of that many steps per launch when
    the cache holds one."""
    e = load_cache(path).get(erosion_key(V))
    if not e:
        return None
    if steps is not None and e.get("steps", 2) != steps:
        e = e.get(f"steps{steps}")
    return _tuning(e) if e else None


def tune_erosion(make_graph, heights=None, steps=20, reps=3, path=CACHE_PATH, persist=True,
                 variants=None, chunks=None, graph_chunks=GRAPH_CHUNKS, rain=1.0,
                 step_options=STEPS):
    """Tune erosion for one mesh.  make_graph(chunk, steps_per_launch)
    returns the mesh's ErosionGraph with a tile plan of that many steps per
    launch cut from that starting chunk (0: default); heights (device
    float64 [V] in the graph's numbering; default: a fixed pseudo-random
    field) start every timed run.  For each step count in step_options, the
    kernel variants (variants[steps], default VARIANTS) are timed on the
    default plan, then the chunks (chunks[steps], default CHUNKS) with the
    best variant; the graph chunks are timed for the best of all.  Returns
    the best ErosionTuning and the table of (us per step, steps, chunk,
    variant, graph chunk) timings; the cache keeps the best and, when it is
    not two-step, the best two-step tuning beside it."""
    import torch
    from . import erosion as _er
    variants = {**VARIANTS, **(variants or {})}
    chunks = {**CHUNKS, **(chunks or {})}
    table = []
    st = None
    stream = None

    def time_it(graph, variant, gchunk):
        graph.variant, graph.graph_chunk = variant, gchunk
        st.load(heights)
        st.run(graph, 4, rain)                      # warm (and capture the graphs)
        best = float("inf")
        for _ in range(reps):
            st.load(heights)
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            st.run(graph, steps, rain)
            b.record(stream)
            b.synchronize()
            best = min(best, a.elapsed_time(b) * 1e3 / steps)
        table.append((round(best, 3), graph.steps if graph.plan is not None else 1,
                      graph.chunk, variant, gchunk))
        return best

    gchunk0 = 32 if 32 in graph_chunks or not graph_chunks else graph_chunks[0]
    per_steps = {}          # steps -> (us, ErosionTuning, graph)
    for K in step_options:
        g = make_graph(0, K)
        if st is None:
            V, dev = g.V, g.device
            if heights is None:
                gen = torch.Generator(device="cpu").manual_seed(20261204)
                heights = (torch.rand(V, generator=gen, dtype=torch.float64) * 18000
                           - 9000).to(dev)
            st = _er.ErosionState(V, dev)
            stream = torch.cuda.current_stream(dev)
        if g.plan is None:      # one step per launch: only the graph chunk matters
            t = time_it(g, 0, gchunk0)
            per_steps.setdefault(1, (t, ErosionTuning(0, 0, gchunk0, K), g))
            break
        if g.steps != K:        # no plan of K steps for this mesh
            continue
        v_best = min(variants[K], key=lambda v: time_it(g, v, gchunk0))
        t_best = min(t for t, k, c, v, _ in table if k == K and v == v_best and c == g.chunk)
        c_best, g_best = 0, g
        for c in chunks[K]:
            gc = make_graph(c, K)
            if gc.plan is None or gc.steps != K:
                continue
            t = time_it(gc, v_best, gchunk0)
            if t < t_best:
                c_best, g_best, t_best = c, gc, t
        per_steps[K] = (t_best, ErosionTuning(c_best, v_best, gchunk0, K), g_best)
    if not per_steps:
        raise ValueError(f"no tile plan of {step_options} steps per launch for this mesh")
    _, best, g_best = min(per_steps.values(), key=lambda r: r[0])
    if graph_chunks:
        best.graph_chunk = min(graph_chunks, key=lambda k: time_it(g_best, best.variant, k))
    if persist:
        entry = {"chunk": best.chunk, "variant": best.variant, "graph_chunk": best.graph_chunk,
                 "steps": best.steps, "us_per_step": min(t for t, *_ in table)}
        if best.steps != 2 and 2 in per_steps:
            t2, b2, _ = per_steps[2]
            entry["steps2"] = {"chunk": b2.chunk, "variant": b2.variant,
                               "graph_chunk": best.graph_chunk, "steps": 2,
                               "us_per_step": round(t2, 3)}
        with _lock:
            cache = load_cache(path)
            key = erosion_key(V)
            cache[key] = merge_erosion_entry(cache.get(key), entry, per_steps)
            save_cache(cache, path)
    return best, table


def merge_erosion_entry(prev, entry, searched):
    """The cache entry to store after a search of the step counts in
    `searched` found `entry`: a cached tuning of a step count not searched
    now that is at least as fast stays first, with the new one beside it
    (a search of two-step plans alone keeps a faster four-step tuning)."""
    if prev and prev.get("steps", 2) not in searched \
            and prev.get("us_per_step", math.inf) <= entry["us_per_step"]:
        return dict(prev, **{f"steps{entry['steps']}": entry})
    return entry
