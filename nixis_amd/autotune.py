"""Autotuner for the noise and erosion launches.

Noise (batched form): tile size x blocks per CU of the persistent launch,
keyed by the batch's shape class (log2 of the total vertex count and the
mean octave count).

Erosion: the tile plan's starting chunk (tile size before halving), the
two-step kernel variant (block size and whether neighbour data stays in
registers) and the steps per captured hipGraph, keyed by the mesh's size
class (log2 of its vertex count).  A coordinate search times the variants
on the default plan, then the chunks with the best variant, then the graph
chunks; none of the choices changes the results.

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
CHUNKS = (512, 640, 704, 768, 832, 896)
GRAPH_CHUNKS = (8, 16, 32, 64)


@dataclass
class ErosionTuning:
    chunk: int = 0          # tile plan starting chunk (0: the library default)
    variant: int = 0        # two-step kernel variant
    graph_chunk: int = 32   # erosion steps per captured graph

    def graph_kwargs(self):
        return {"chunk": self.chunk, "variant": self.variant, "graph_chunk": self.graph_chunk}


def erosion_key(V):
    return f"erosion:V2^{int(math.log2(max(int(V), 1)))}"


def lookup_erosion(V, path=CACHE_PATH):
    """Cached erosion tuning for meshes of V vertices' size class, or None."""
    e = load_cache(path).get(erosion_key(V))
    return ErosionTuning(e["chunk"], e["variant"], e["graph_chunk"]) if e else None


def tune_erosion(make_graph, heights=None, steps=20, reps=3, path=CACHE_PATH, persist=True,
                 variants=PAIR_VARIANTS, chunks=CHUNKS, graph_chunks=GRAPH_CHUNKS, rain=1.0):
    """Tune erosion for one mesh.  make_graph(chunk) returns the mesh's
    ErosionGraph with a tile plan cut from that starting chunk (0: default);
    heights (device float64 [V] in the graph's numbering; default: a fixed
    pseudo-random field) start every timed run.  Returns the best
    ErosionTuning and the table of (us per step, chunk, variant, graph
    chunk) timings."""
    import torch
    from . import erosion as _er
    g = make_graph(0)
    V = g.V
    dev = g.device
    if heights is None:
        gen = torch.Generator(device="cpu").manual_seed(20261204)
        heights = (torch.rand(V, generator=gen, dtype=torch.float64) * 18000 - 9000).to(dev)
    st = _er.ErosionState(V, dev)
    stream = torch.cuda.current_stream(dev)
    table = []

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
        table.append((round(best, 3), graph.chunk, variant, gchunk))
        return best

    if g.plan is None:      # one step per launch: only the graph chunk matters
        variants, chunks = (0,), ()
    gchunk0 = 32 if 32 in graph_chunks or not graph_chunks else graph_chunks[0]
    v_best = min(variants, key=lambda v: time_it(g, v, gchunk0))
    c_best, g_best = 0, g
    t_best = min(t for t, c, v, _ in table if v == v_best and c == g.chunk)
    for c in chunks:
        gc = make_graph(c)
        if gc.plan is None:
            continue
        t = time_it(gc, v_best, gchunk0)
        if t < t_best:
            c_best, g_best, t_best = c, gc, t
    gc_best = min(graph_chunks, key=lambda k: time_it(g_best, v_best, k)) if graph_chunks else gchunk0
    best = ErosionTuning(c_best, v_best, gc_best)
    if persist:
        with _lock:
            cache = load_cache(path)
            cache[erosion_key(V)] = {"chunk": best.chunk, "variant": best.variant,
                                     "graph_chunk": best.graph_chunk,
                                     "us_per_step": min(t for t, *_ in table)}
            save_cache(cache, path)
    return best, table
