"""Autotuner for the batched noise launch (tile size x blocks per CU).

Candidates are timed with HIP events on the launch stream; the best one is
stored in a JSON file keyed by the batch's shape class (log2 of the total
vertex count and the mean octave count), so later batches of the same class
reuse it without timing.
"""
import json
import math
import os
import threading

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
