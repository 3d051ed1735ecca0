#!/usr/bin/env python3
"""Benchmark of the terrain hot path on one MI355X.

Default workload (the headline `value`): fused OpenSimplex noise at scale --
d=2000 icosphere (40,000,002 vertices), 16 octaves, base frequency 2.0,
lacunarity 2.0, persistence 0.55, seed 7.  One step = one fused pass of all
16 octaves over the planet (640,000,032 vertex-octaves), inputs resident in
HBM.  Reported in G vertex-octaves/s.

Also measured in the same run (secondary objects of the JSON line):
  erosion  -- d=1000 icosphere (10,000,002 vertices), 12-octave seed-99
              terrain rescaled to [-9000, 9000], rain 1.0 per vertex per
              iteration, 500 iterations (G vertex-iterations/s);
  cpu_baseline -- the FP64 oracle (oracle/nx_oracle.c, OpenMP) on the
              host cores on the same noise workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-erosion]
       [--no-cpu] [--erosion-iters N]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gvertex-octaves/s noise; Gvertex-iters/s erosion; p50/p99 planet latency"
# Algorithmic work per unit (worst-case octahedral region, all 8 lattice
# points active): 31 FP64 + 169 FP32 flops per vertex-octave; MI355X vector
# peaks 78.6 TF/s FP64 and 157.3 TF/s FP32.
NOISE_FLOPS = 200
NOISE_T_MIN = 31 / 78.6e12 + 169 / 157.3e12          # s per vertex-octave
NOISE_PEAK_TFLOPS = NOISE_FLOPS / NOISE_T_MIN / 1e12  # blended peak, 136.2
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-erosion", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--erosion-iters", type=int, default=500)
    ap.add_argument("--noise-d", type=int, default=2000)
    ap.add_argument("--erosion-d", type=int, default=1000)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-export", action="store_true")
    ap.add_argument("--no-climate", action="store_true")
    ap.add_argument("--export-d", type=int, default=2500)
    ap.add_argument("--served-jobs", type=int, default=120)
    ap.add_argument("--served-load", type=float, default=0.8,
                    help="arrival rate as a fraction of the plain form's capacity")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three plain 1 GiB device copies (for scaling "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def profiled_traffic(kind, V, steps_per_launch=None):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (profiles/traffic.json, copy-
    calibrated), when it was measured on this workload size (and, for
    erosion, the same steps per launch); else None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            e = json.load(f).get(kind)
    except (OSError, ValueError):
        return None
    if not e or e.get("V") != V:
        return None
    if steps_per_launch is not None and e.get("steps_per_launch") != steps_per_launch:
        return None
    return int(e["bytes_per_launch"])


def calibrate_copy():
    """Three device-to-device copies of 1 GiB: a kernel whose HBM traffic is
    known (1 GiB read + 1 GiB written per launch)."""
    import torch
    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 4, "launches": 3}), flush=True)


def bench_noise(args):
    import torch
    from nixis_amd import opensimplex as osx, terrain, util
    d, n_oct, f0, lac, pers, seed = args.noise_d, 16, 2.0, 2.0, 0.55, 7
    t0 = time.perf_counter()
    pts = util.create_points(d)
    V = pts.shape[0]
    log(f"[noise] mesh d={d} V={V} built in {time.perf_counter() - t0:.1f}s")
    perm, pgi = osx.init(seed)
    rough, strength = terrain.octave_scalars(n_oct, f0, 0.4, lac, pers, 1.0)
    dv = torch.from_numpy(pts).cuda()
    out = torch.empty(V, dtype=torch.float64, device="cuda")
    for _ in range(args.warmup):
        terrain.sample_octaves_device(dv, None, perm, pgi, rough, strength, 1.0, out=out)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        terrain.sample_octaves_device(dv, None, perm, pgi, rough, strength, 1.0, out=out)
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    ms = start.elapsed_time(end) / args.steps
    vo = V * n_oct
    gvo = vo / (ms * 1e-3) / 1e9
    log(f"[noise] {ms:.3f} ms/step (events), wall {wall / args.steps * 1e3:.3f} ms/step, "
        f"{gvo:.1f} Gvo/s")
    return dict(V=V, vo=vo, ms=ms, gvo=gvo, pts=pts, perm=perm, pgi=pgi, n_oct=n_oct, f0=f0,
                lac=lac, pers=pers, out=out.cpu().numpy())


def cpu_noise_baseline(nz, budget_s=25.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    pts = nz["pts"]
    # warm (first call loads the library and spins up the thread pool)
    O.sample_octaves(pts[:10000], None, nz["perm"], nz["pgi"], nz["n_oct"], nz["f0"], 0.4,
                     nz["lac"], nz["pers"], 1.0)
    times = []
    t_all = time.perf_counter()
    while len(times) < 5 and (time.perf_counter() - t_all) < budget_s:
        t = time.perf_counter()
        O.sample_octaves(pts, None, nz["perm"], nz["pgi"], nz["n_oct"], nz["f0"], 0.4, nz["lac"],
                         nz["pers"], 1.0)
        times.append(time.perf_counter() - t)
    med = statistics.median(times)
    val = nz["vo"] / med / 1e9
    # the last oracle run's elevations against the GPU's on the same planet
    ref = O.sample_octaves(pts[::101], None, nz["perm"], nz["pgi"], nz["n_oct"], nz["f0"], 0.4,
                           nz["lac"], nz["pers"], 1.0)
    gpu = nz["out"][::101]
    bound = 0.5 * 0.4 * sum(nz["pers"] ** k for k in range(nz["n_oct"])) * 1e-5
    # per-octave noise and lattice selection of the same vertices: a debug
    # launch of the GPU kernel against the oracle's per-octave values
    import torch
    from nixis_amd import _lib, terrain
    sub = np.ascontiguousarray(pts[::101])
    rough, strength = terrain.octave_scalars(nz["n_oct"], nz["f0"], 0.4, nz["lac"], nz["pers"], 1.0)
    dv = torch.from_numpy(sub).to("cuda")
    dn = torch.empty((nz["n_oct"], sub.shape[0]), dtype=torch.float32, device="cuda")
    ds = torch.empty((nz["n_oct"], sub.shape[0], 5), dtype=torch.int32, device="cuda")
    terrain.sample_octaves_device(dv, None, nz["perm"], nz["pgi"], rough, strength, 1.0, dn, ds)
    _lib.Context.get(dv.device).sync_check()
    _, raw, sel = O.sample_octaves(sub, None, nz["perm"], nz["pgi"], nz["n_oct"], nz["f0"], 0.4,
                                   nz["lac"], nz["pers"], 1.0, raw=True, sel=True)
    per_oct = np.abs(dn.cpu().numpy().astype(np.float64) - raw).max(axis=1)
    mism = int((ds.cpu().numpy() != sel).any(axis=2).sum())
    parity = {"vertices_checked": int(ref.size), "max_abs_elev_err": float(np.abs(gpu - ref).max()),
              "bound": bound, "per_octave_max_abs_noise_err": [float(f"{x:.3e}") for x in per_oct],
              "noise_tolerance": 1e-5, "selection_mismatches": mism}
    log(f"[cpu] per-octave max |noise - fp64| over {sub.shape[0]} vertices: "
        + " ".join(f"{x:.1e}" for x in per_oct) + f"; selection mismatches {mism}")
    log(f"[cpu] oracle noise {val:.3f} Gvo/s over {len(times)} runs, {O.num_threads()} threads")
    return {"value": round(val, 4), "unit": "Gvertex-octaves/s", "cores": O.num_threads(),
            "kind": "port",
            "sample": f"full noise workload ({nz['V']} vertices x {nz['n_oct']} octaves), "
                      f"median of {len(times)} warm runs, FP64 oracle with OpenMP",
            "parity_spot_check": parity}


def bench_erosion(args):
    import numpy as np
    import torch
    from nixis_amd import erosion, opensimplex as osx, terrain, util
    d, iters = args.erosion_d, args.erosion_iters
    t0 = time.perf_counter()
    pts, nb = util.icosphere_with_neighbors(d)
    V = pts.shape[0]
    log(f"[erosion] mesh+adjacency d={d} V={V} in {time.perf_counter() - t0:.1f}s")
    perm, pgi = osx.init(99)
    rough, strength = terrain.octave_scalars(12, 1.5, 0.4, 2.0, 0.5, 1.0)
    dv = torch.from_numpy(pts).cuda()
    h = terrain.sample_octaves_device(dv, None, perm, pgi, rough, strength, 1.0)
    h, _ = util.rescale_device(h, -9000, 9000)
    from nixis_amd import autotune
    t0 = time.perf_counter()
    order = util.sfc_order(pts)
    graph = erosion.ErosionGraph(pts, nb, order=order)      # SFC order + 4-step tile plan
    hd = graph.permute_device(h)
    log(f"[erosion] SFC order, ELL table and tile plan in {time.perf_counter() - t0:.1f}s")
    # steps per launch, tile chunk, kernel variant and graph chunk from the autotuner
    tuning = autotune.lookup_erosion(V)
    tuned_now = tuning is None
    if tuned_now:
        t0 = time.perf_counter()
        tuning, table = autotune.tune_erosion(
            lambda c, k: graph if (c, k) == (0, graph.steps)
            else erosion.ErosionGraph(pts, nb, order=order, chunk=c, steps=k),
            heights=hd, steps=20, reps=2)
        log(f"[erosion] autotuned in {time.perf_counter() - t0:.1f}s: {tuning.steps} steps per "
            f"launch, chunk {tuning.chunk} variant {tuning.variant} graph chunk "
            f"{tuning.graph_chunk}; {len(table)} candidates, "
            f"{min(table)[0]:.1f}..{max(table)[0]:.1f} us/step")
    if (tuning.chunk, tuning.steps) != (graph.chunk, graph.steps):
        graph = erosion.ErosionGraph(pts, nb, order=order, **tuning.graph_kwargs())
    else:
        graph.variant, graph.graph_chunk = tuning.variant, tuning.graph_chunk
    st = erosion.ErosionState(V)
    nnz = int((nb >= 0).sum())
    plan = graph.plan
    log("[erosion] " + (f"{plan.ntiles} tiles" if plan is not None else "no tile plan"))
    # warm-up run, then the timed run from the same initial state
    st.load(hd)
    st.run(graph, min(args.warmup * 10, iters), 1.0)
    st.load(hd)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    st.run(graph, iters, 1.0)
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end)
    hh = torch.empty(V, dtype=torch.float64, device="cuda")
    ss = torch.empty(V, dtype=torch.float64, device="cuda")
    st.store(hh, None, ss)
    st.ctx.sync_check()
    # erosion.py:262-267 moves height into sediment and back, so h + s stays
    # each vertex's starting height in exact arithmetic; in FP64 each step
    # rounds it by at most about one ulp of |h| + |s|.  Drift in those ulps:
    eps = 2.0 ** -52
    scale = (hh.abs() + ss.abs() + hd.abs()) * eps
    drift = ((hh + ss) - hd).abs()
    ulps = torch.where(scale > 0, drift / scale, drift)
    finite = torch.isfinite(ulps)
    drift_ulps_max = float(ulps[finite].max()) if bool(finite.any()) else float("nan")
    del ss, scale, drift, ulps, finite
    hh = graph.unpermute_device(hh).cpu().numpy()
    gvi = V * iters / (ms * 1e-3) / 1e9
    B = 28 * V + 8 * nnz
    per_step_ms = ms / iters
    spl = plan.steps if plan is not None else 1
    launch_us = spl * per_step_ms * 1e3
    traffic = profiled_traffic("erosion", V, spl)
    gbs = B / (per_step_ms * 1e-3) / 1e9
    log(f"[erosion] {iters} iters in {ms:.2f} ms ({per_step_ms * 1e3:.1f} us/iter), "
        f"{gvi:.2f} Gvi/s, {gbs:.0f} GB/s algorithmic")
    return {
        "workload": f"erosion at scale: d={d} icosphere ({V} vertices, {nnz} neighbour slots), "
                    f"12-octave seed-99 terrain rescaled to [-9000, 9000], rain 1.0, "
                    f"{iters} iterations, FP64 fields in the reference's operation order",
        "value": round(gvi, 3), "unit": "Gvertex-iters/s", "ms_total": round(ms, 3),
        "us_per_iter": round(per_step_ms * 1e3, 2),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "algorithmic_bytes_per_launch": spl * B, "steps_per_launch": spl,
                     "plan_stream_bytes_per_launch": getattr(graph, "stream_bytes_per_launch", None),
                     "launch_us": round(launch_us, 2),
                     # B counts a pass over fields and graph per step; a launch of
                     # several steps makes one, so `frac` can pass 1.  The rate of
                     # the bytes the launch does move:
                     "traffic_GBs": round(traffic / (launch_us * 1e3), 1) if traffic else None,
                     "traffic_frac": (round(traffic / (launch_us * 1e3) / HBM_PEAK_GBS, 4)
                                      if traffic else None)},
        "tuning": {"steps_per_launch": spl, "chunk": tuning.chunk, "variant": tuning.variant,
                   "graph_chunk": tuning.graph_chunk,
                   "tiles": plan.ntiles if plan is not None else 0,
                   "source": "autotuned now" if tuned_now else "autotune cache"},
        "checks": {"finite": bool(np.isfinite(hh).all()),
                   "h_plus_s_drift_max_ulps": round(drift_ulps_max, 2),
                   "h_plus_s_drift_bound_ulps": iters,
                   "height_range": float(hh.max() - hh.min())},
        "_graph": graph, "_h": h, "_V": V,
    }


# min/max, summary and writing passes: reads 8+8+8, writes 8+1
CHAIN_BYTES_PER_VERTEX = 33
CHAIN5_BYTES_PER_VERTEX = 75  # the five-pass form: reads 8+8+9+9+8, writes 9+8+8+8


def bench_post_chain(args, nz):
    """nixis.py:330-364 on the noise output (d=2000): rescale to
    [-4000, 8850], ocean level at 55%, mask, power_rescale 0.5 / 2.0, shift,
    split rescale.  Each timed call gets a fresh copy of the noise heights
    (copied outside the timed events)."""
    import torch
    from nixis_amd import _lib, util
    src = torch.from_numpy(nz["out"]).cuda()
    h = torch.empty_like(src)
    n = src.numel()
    L = _lib.lib()
    ws = torch.empty(L.nx_height_chain_workspace_bytes(n), dtype=torch.uint8, device="cuda")
    for _ in range(args.warmup):
        h.copy_(src)
        util.height_chain_device(h, -4000.0, 8850.0, 55.0, workspace=ws)
    total = 0.0
    for _ in range(args.steps):
        h.copy_(src)
        total += _events_ms(lambda: util.height_chain_device(h, -4000.0, 8850.0, 55.0,
                                                             workspace=ws))
    ctx = _lib.Context.get(h.device)
    ctx.sync_check()
    ms = total / args.steps
    gbs = CHAIN_BYTES_PER_VERTEX * n / (ms * 1e-3) / 1e9
    log(f"[chain] {ms:.3f} ms per chain over {n} heights, {gbs:.0f} GB/s")
    # the five-pass form for comparison (same results)
    ctx.set_chain_passes(5)
    try:
        t5 = 0.0
        for i in range(args.warmup + args.steps):
            h.copy_(src)
            t = _events_ms(lambda: util.height_chain_device(h, -4000.0, 8850.0, 55.0,
                                                            workspace=ws))
            t5 += t if i >= args.warmup else 0.0
        ctx.sync_check()
    finally:
        ctx.set_chain_passes(2)
    ms5 = t5 / args.steps
    log(f"[chain] five-pass form: {ms5:.3f} ms, "
        f"{CHAIN5_BYTES_PER_VERTEX * n / (ms5 * 1e-3) / 1e9:.0f} GB/s")
    # noise and chain as one call (min/max reduced in the noise kernel)
    from nixis_amd import terrain
    rough, strength = terrain.octave_scalars(nz["n_oct"], nz["f0"], 0.4, nz["lac"], nz["pers"],
                                             1.0)
    dv = torch.from_numpy(nz["pts"]).cuda()

    def fused():
        terrain.terrain_heights_device(dv, nz["perm"], nz["pgi"], rough, strength, 1.0, -4000.0,
                                       8850.0, 55.0, workspace=ws, out=h)
    for _ in range(args.warmup):
        fused()
    fused_ms = _events_ms(lambda: [fused() for _ in range(args.steps)]) / args.steps
    _lib.Context.get(h.device).sync_check()
    log(f"[chain] noise + chain in one call: {fused_ms:.3f} ms "
        f"(noise alone {nz['ms']:.3f} ms + chain {ms:.3f} ms)")
    return {"workload": f"height chain nixis.py:330-364 on the d={args.noise_d} noise heights "
                        f"({n} vertices): rescale [-4000, 8850], ocean 55%, power 0.5 / 2.0, "
                        "shift, split rescale",
            "ms": round(ms, 4), "Gvertex_per_s": round(n / (ms * 1e-3) / 1e9, 2),
            "five_pass_ms": round(ms5, 4),
            "fused_noise_chain_ms": round(fused_ms, 4), "noise_ms": round(nz["ms"], 4),
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "bytes_per_vertex": CHAIN_BYTES_PER_VERTEX,
                         "traffic": profiled_traffic("chain", n)}}


def bench_export(args):
    """nixis.py:270-302 and 516-540 on the device: the pixel query of a
    d-division planet (grid build = the reference's KD-tree build, then the
    3 nearest vertices of every pixel of make_ll_arr) and the gray sampling
    of a height layer, for the reference's default 2048 x 1024 image and a
    16384 x 8192 one.  The mesh is generated on the device."""
    import torch
    from nixis_amd import _lib, util
    d = args.export_d
    pts, _ = util.create_mesh_device(d, cells=False)
    V = pts.shape[0]
    colors = torch.rand(V, dtype=torch.float64, device="cuda") * 255.0
    util.KnnGrid(pts)                                       # warm-up build
    torch.cuda.synchronize()
    t = time.perf_counter()
    grid = util.KnnGrid(pts)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t) * 1e3
    out = {"workload": f"d={d} icosphere ({V} vertices) made on the device; grid build, then "
                       "make_ll_arr + 3-nearest query and gray sampling per image",
           "grid_build_ms": round(build_ms, 2), "images": {}}
    for (w, h) in ((2048, 1024), (16384, 8192)):
        grid.query_pixels_device(w, h, 1.0, 3)              # warm-up
        dd, nn = grid.query_pixels_device(w, h, 1.0, 3)
        q_ms = _events_ms(lambda: grid.query_pixels_device(w, h, 1.0, 3))
        util.make_gray_array_device(dd, nn, colors)
        g_ms = _events_ms(lambda: util.make_gray_array_device(dd, nn, colors))
        _lib.Context.get(pts.device).sync_check()
        npix = w * h
        out["images"][f"{w}x{h}"] = {"query_ms": round(q_ms, 3), "gray_ms": round(g_ms, 3),
                                      "Mpixel_per_s_query": round(npix / (q_ms * 1e-3) / 1e6, 1)}
        log(f"[export] d={d} V={V}: {w}x{h} query {q_ms:.2f} ms, gray {g_ms:.2f} ms")
        del dd, nn
    log(f"[export] grid build {build_ms:.1f} ms")
    return out


def bench_climate(args):
    """climate.py:351-590 on the device for a d=1000 planet (10,000,002
    vertices): a year of insolation (360 daily latitude tables, each 181
    latitudes x 360 rotation steps, then 360 interpolations per vertex), one
    day of 360 rotation steps per vertex (brute_daily_insolation) and the
    starting surface temperatures (assign_surface_temp)."""
    import torch
    from nixis_amd import climate, util
    pts, _ = util.create_mesh_device(1000, cells=False)
    V = pts.shape[0]
    ws = climate._ws(pts.device)
    out = torch.empty(V, dtype=torch.float32, device=pts.device)
    climate.yearly_insolation_device(pts, 1.0, 23.44, out=out, workspace=ws)
    y_ms = _events_ms(lambda: climate.yearly_insolation_device(pts, 1.0, 23.44, out=out,
                                                              workspace=ws))
    out.zero_()
    climate.sample_insolation_device(out, pts, 1.0, -180.0, 23.44, 360, 1.0, workspace=ws)
    out.zero_()
    d_ms = _events_ms(lambda: climate.sample_insolation_device(out, pts, 1.0, -180.0, 23.44, 360,
                                                               1.0, workspace=ws))
    t_ms = _events_ms(lambda: climate.assign_surface_temp_device(pts, 1.0, 23.44, out=out))
    log(f"[climate] V={V}: year {y_ms:.2f} ms, day of 360 rotations {d_ms:.2f} ms, "
        f"surface temperatures {t_ms:.3f} ms")
    return {"workload": f"insolation on a d=1000 icosphere ({V} vertices), tilt 23.44",
            "yearly_ms": round(y_ms, 3), "daily_360_rotations_ms": round(d_ms, 3),
            "Gvertex_rotations_per_s": round(V * 360 / (d_ms * 1e-3) / 1e9, 2),
            "surface_temp_ms": round(t_ms, 4)}


def cpu_erosion_baseline(er, iters=4):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from nixis_amd import util
    d = int(round(((er["_V"] - 2) / 10) ** 0.5))
    pts, nb = util.icosphere_with_neighbors(d)
    import numpy as np
    from nixis_amd import erosion
    h = er["_h"].cpu().numpy()
    href = h.copy()
    t = time.perf_counter()
    O.erode_terrain3(pts, nb, href, iters, 1.0)
    dt = time.perf_counter() - t
    val = er["_V"] * iters / dt / 1e9
    # the same iterations through the GPU drop-in, against the oracle's heights
    hg = h.copy()
    erosion.erode_terrain3(pts, nb, hg, iters, rain=1.0)
    span = float(href.max() - href.min())
    parity = {"vertices": int(h.size), "iterations": iters,
              "mismatches": int((hg != href).sum()),
              "max_abs_err_over_range": float(np.abs(hg - href).max() / span)}
    log(f"[cpu] oracle erosion {val:.4f} Gvi/s; GPU drop-in after {iters} iterations: "
        f"{parity['mismatches']} of {h.size} heights differ from the oracle")
    return {"value": round(val, 5), "unit": "Gvertex-iters/s", "cores": O.num_threads(),
            "kind": "port", "sample": f"{iters} of the {er['_V']}-vertex erosion iterations, "
                                      "FP64 oracle with OpenMP",
            "parity_check": parity}


def _events_ms(fn, stream=None):
    import torch
    s = stream or torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(s)
    fn()
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b)


def bench_batched(args):
    """64 planets, d log-uniform in [64, 1024], seeds 0..63, octaves 4..16,
    0..200 erosion steps: one batched run (persistent noise launch, batched
    rescale/load, one graph-captured launch per erosion step) against the
    same jobs one planet per call."""
    from nixis_amd import autotune, batch, terrain
    jobs = batch.load_jobs(batch.BATCHED_64_PATH)
    t0 = time.perf_counter()
    meshes = batch.MeshCache()
    runs = [batch.PlanetRun(j, meshes) for j in jobs]
    log(f"[batched] {len(jobs)} planets, {sum(j.V for j in jobs)} vertices, meshes in "
        f"{time.perf_counter() - t0:.1f}s")
    br = batch.BatchRunner()
    tuning = autotune.lookup(runs)
    tuned_now = tuning is None
    if tuned_now:
        tuning, table = autotune.tune(br, runs, reps=2)
        log(f"[batched] autotuned tile={tuning.tile} blocks/CU={tuning.blocks_per_cu}")
    br.run(runs, tuning)                       # warm (captures the step graph)
    batch.run_plain(runs)
    ms_b = min(_events_ms(lambda: br.run(runs, tuning)) for _ in range(2))
    ms_noise = _events_ms(lambda: br.noise(runs, tuning))
    ms_p = min(_events_ms(lambda: batch.run_plain(runs)) for _ in range(2))
    def plain_noise():
        # the batched noise phase's work one planet per call: noise, the
        # rescale and the erosion state load
        from nixis_amd import util
        for r in runs:
            j = r.job
            terrain.sample_octaves_device(r.verts, None, r.perm, r.pgi, r.rough, r.strength,
                                          j.world_radius, out=r.elev)
            if j.lower < j.upper:
                util.rescale_device(r.elev, j.lower, j.upper, out=r.elev)
            if r.state is not None:
                r.state.load(r.elev)
    ms_pnoise = _events_ms(plain_noise)
    vo = sum(j.vertex_octaves for j in jobs)
    vi = sum(j.vertex_iters for j in jobs)
    log(f"[batched] batched {ms_b:.1f} ms (noise {ms_noise:.2f} ms), plain {ms_p:.1f} ms "
        f"(noise {ms_pnoise:.2f} ms)")
    return {
        "workload": "64 planets: d log-uniform in [64,1024] (even), seeds 0..63, octaves "
                    "U{4..16}, erosion steps U{0..200}, rescale to [-9000,9000], rain 0.3/320; "
                    "job list nixis_amd/jobs/batched_64planets.json (numpy default_rng(20261202))",
        "planets": len(jobs), "vertices": sum(j.V for j in jobs),
        "vertex_octaves": vo, "vertex_iters": vi,
        "batched_ms": round(ms_b, 3), "plain_ms": round(ms_p, 3),
        "speedup_vs_plain": round(ms_p / ms_b, 3),
        "noise_batched_ms": round(ms_noise, 3), "noise_plain_ms": round(ms_pnoise, 3),
        "Gvertex_octaves_per_s_noise_phase": round(vo / (ms_noise * 1e-3) / 1e9, 2),
        "Gvertex_iters_per_s_whole_batch": round(vi / (ms_b * 1e-3) / 1e9, 2),
        "tuning": {"tile": tuning.tile, "blocks_per_cu": tuning.blocks_per_cu,
                   "source": "autotuned now" if tuned_now else "autotune cache"},
    }


def bench_served(args):
    """Poisson arrivals of planet jobs (d log-uniform in [32,2048], octaves
    4..16, 30% with 100..300 erosion steps) served with continuous batching
    under an HBM budget, against the plain one-planet-per-call form on the
    same trace.  The arrival rate is set to `served_load` times the plain
    form's capacity, measured by running the jobs back to back."""
    from nixis_amd import batch, serve
    n = args.served_jobs

    def trace(rate):
        if n == 120:
            return serve.load_trace(batch.SERVED_120_PATH, rate)
        return serve.served_jobs(n, rate=rate)
    proto = trace(1.0)
    t0 = time.perf_counter()
    log(f"[served] preparing meshes for {len(set(j.d for j in proto))} subdivisions")
    plain = serve.PlainServer()
    served = serve.PlanetServer(hbm_budget=64 << 30)
    serve._prepare_meshes(plain.meshes, proto)
    served.meshes = plain.meshes   # same per-subdivision assets for both
    log(f"[served] {n} jobs, meshes in {time.perf_counter() - t0:.1f}s")
    # capacity of the plain form: all jobs back to back
    b2b = [batch.Job(**{**j.__dict__, "arrival": 0.0}) for j in proto]
    plain.serve(b2b[: min(8, n)], prepare_meshes=False)       # warm
    outs = plain.serve(b2b, prepare_meshes=False)
    busy = max(o.done for o in outs)
    rate = args.served_load * n / busy
    jobs = trace(rate)
    log(f"[served] plain back-to-back busy {busy:.2f}s -> rate {rate:.1f} jobs/s")
    out_plain = plain.serve(jobs, prepare_meshes=False)
    out_served = served.serve(jobs, prepare_meshes=False)
    res_plain = serve.latency_summary(out_plain)
    res_served = serve.latency_summary(out_served)
    log(f"[served] served p50 {res_served['p50_ms']:.1f} ms p99 {res_served['p99_ms']:.1f} ms; "
        f"plain p50 {res_plain['p50_ms']:.1f} ms p99 {res_plain['p99_ms']:.1f} ms")
    for name, outs in (("served", out_served), ("plain", out_plain)):
        worst = sorted(outs, key=lambda o: -o.latency)[:4]
        log(f"[served] {name} slowest: " + ", ".join(
            f"d={o.job.d} it={o.job.erosion_iters} arr={o.job.arrival:.3f}s lat={o.latency * 1e3:.0f}ms"
            for o in worst))
    # capacity: every job waiting at t=0, both forms
    cap_s = serve.latency_summary(served.serve(b2b, prepare_meshes=False))
    log(f"[served] all jobs at t=0: served makespan {cap_s['makespan_s']:.3f}s, "
        f"plain {busy:.3f}s")
    return {
        "workload": f"{n} jobs, Poisson arrivals at {rate:.2f} jobs/s ({args.served_load:.0%} of "
                    "the plain form's back-to-back capacity), d log-uniform in [32,2048], "
                    "octaves U{4..16}, 30% with U{100..300} erosion steps; trace "
                    "nixis_amd/jobs/served_120jobs_rate1.json (numpy default_rng(20261203))",
        "rate_jobs_per_s": round(rate, 3), "hbm_budget_GB": 64,
        "served": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res_served.items()},
        "plain": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res_plain.items()},
        "peak_job_bytes": int(served.peak_bytes),
        "capacity_all_at_t0": {"served_makespan_s": round(cap_s["makespan_s"], 4),
                               "plain_makespan_s": round(busy, 4),
                               "served_jobs_per_s": round(n / cap_s["makespan_s"], 2),
                               "plain_jobs_per_s": round(n / busy, 2)},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the terrain path "
              "shards whole planets per GPU and is run here on a single MI355X.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    nz = bench_noise(args)
    achieved = NOISE_FLOPS * nz["gvo"] * 1e9 / 1e12
    line = {
        "metric": METRIC,
        "value": round(nz["gvo"], 3),
        "unit": "Gvertex-octaves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(nz["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {
            "workload": f"noise at scale: d={args.noise_d} icosphere ({nz['V']} vertices), "
                        f"16 octaves, f0 2.0, lacunarity 2.0, persistence 0.55, seed 7; "
                        "FP64 skew/floor/hash, FP32 gradients",
            "vertex_octaves_per_step": nz["vo"],
        },
        "roofline": {"bound": "valu", "achieved": round(achieved, 2),
                     "peak": round(NOISE_PEAK_TFLOPS, 1), "unit": "TFLOP/s",
                     "frac": round(achieved / NOISE_PEAK_TFLOPS, 4),
                     "traffic": profiled_traffic("noise", nz["V"]),
                     "flops_per_vertex_octave": NOISE_FLOPS},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_noise_baseline(nz)
    line["post_chain"] = bench_post_chain(args, nz)
    if not args.no_erosion:
        er = bench_erosion(args)
        if not args.no_cpu:
            er["cpu_baseline"] = cpu_erosion_baseline(er)
        for k in [k for k in er if k.startswith("_")]:
            er.pop(k)
        line["erosion"] = er
    if not args.no_batched:
        line["batched"] = bench_batched(args)
    if not args.no_served:
        line["served"] = bench_served(args)
    if not args.no_export:
        line["export"] = bench_export(args)
    if not args.no_climate:
        line["climate"] = bench_climate(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
