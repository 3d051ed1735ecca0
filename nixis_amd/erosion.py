"""Drop-in for the reference's erosion.py (version 3 of the erosion step).

``erode_terrain3`` (erosion.py:171-192) and ``erosion_iteration3``
(erosion.py:197-279) keep their signatures and in-place semantics; the
Jacobi steps run on the GPU in FP64, operation for operation as the
reference computes them (see nixis_amd/csrc/nx_erosion.hip).  ``ErosionGraph`` / ``ErosionState`` are the
device-resident forms used by the batched and served paths.
"""
import ctypes
import math

import numpy as np

from . import _lib

RAIN_DEFAULT = 0.3 / 320  # erosion.py:182


def calc_slope(v0, v1, dist):
    """erosion.py:34-36."""
    return (v1 - v0) / (dist + 0.00001)


def calc_distance(v0, v1):
    """erosion.py:38-40."""
    return math.sqrt((v0[0] - v1[0]) ** 2 + (v0[1] - v1[1]) ** 2 + (v0[2] - v1[2]) ** 2)


def _plan_stream_bytes(plan, img, V):
    """Bytes one launch over this plan reads as contiguous runs from HBM:
    every vertex's h, w and s (24 B), the slot codes (6 x 2 B) and ring
    indices (4 B) of each tile's first-step and loaded vertices, and its
    edge lengths (8 B each); the gathers of ring fields come on top and may
    hit in L2.  Summed from the plan image's per-tile tables."""
    nt = int(plan.ntiles)
    K = int(plan.steps)

    def tab(off, dt, n):
        return np.frombuffer(img, dtype=dt, count=n, offset=int(off)).astype(np.int64)

    T = np.diff(tab(plan.off_start, np.int32, nt + 1))
    if K == 2:
        cs, load = tab(plan.off_cs, np.int32, nt), tab(plan.off_load, np.int32, nt)
    else:
        pref = tab(plan.off_pref, np.int32, nt * (K + 1)).reshape(nt, K + 1)
        cs, load = pref[:, K - 1], pref[:, K]
    lengths = int(tab(plan.off_ebase, np.int64, nt + 1)[-1])
    return 24 * int(V) + 12 * int(cs.sum()) + 8 * lengths + 4 * int((load - T).sum())


class ErosionGraph:
    """Device neighbour table: col int32 [6,V] and FP64 elen [6,V],
    plus the tile plan that lets erosion run several steps per launch.

    Built on the host from a reference-style neighbour table ([V,6] int32,
    -1 padded) and FP64 vertex coordinates.  `order` renumbers the vertices
    on the device: "sfc" (default, 3D Morton order: compact tiles), None
    (caller order) or an explicit order[new] = old.  Fields passed to
    ErosionState use the device numbering (``permute``/``unpermute``).
    With `tiled`, the tile plan is built when the graph cuts into tiles
    (nx_tile_plan_build_chunk); otherwise, or for graphs without locality,
    every step is its own launch.  `steps` (4 or 2) is the number of steps
    one launch covers: a 4-step plan carries four rings per tile (a graph
    whose tiles cannot hold them gets a 2-step plan; ``steps`` tells which).
    `chunk` is the plan's starting tile size (0: the library's default),
    `variant` the kernel variant for the plan's step count (the context's
    NX_OPT_MULTI_KERNEL or NX_OPT_PAIR_KERNEL) and `graph_chunk` the steps
    per captured graph that runs on this graph use (None: the context's
    setting); autotune.tune_erosion picks the four per mesh size.  None of
    them changes the results.
    """

    def __init__(self, verts, neighbors, order="sfc", device=None, tiled=True, steps=4,
                 chunk=0, variant=None, graph_chunk=None):
        import torch
        verts = np.ascontiguousarray(verts, dtype=np.float64)
        nb = np.ascontiguousarray(neighbors, dtype=np.int32)
        if verts.ndim != 2 or verts.shape[1] != 3 or nb.ndim != 2 or nb.shape[1] != 6 \
                or nb.shape[0] != verts.shape[0]:
            raise ValueError("verts must be [V,3] float64 and neighbors [V,6] int32")
        V = nb.shape[0]
        if isinstance(order, str):
            if order != "sfc":
                raise ValueError(f"unknown order {order!r}")
            from . import util
            order = util.sfc_order(verts)
        col = np.empty((6, V), dtype=np.int32)
        elen = np.empty((6, V), dtype=np.float64)
        o = None if order is None else np.ascontiguousarray(order, dtype=np.int64)
        _lib.check(_lib.lib().nx_ell_from_adjacency(_lib.ptr(nb), V, _lib.ptr(verts),
                                                    _lib.ptr(o), _lib.ptr(col), _lib.ptr(elen)))
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.V = V
        self.order = o
        self.col = torch.from_numpy(col).to(device)
        self.elen = torch.from_numpy(elen).to(device)
        self.device = self.col.device
        self.plan = None
        self.stream_bytes_per_launch = None
        self._plan_blob = None
        self.steps = int(steps)
        self._tuning(chunk, variant, graph_chunk)
        if tiled:
            self._build_plan(col, elen)

    def _tuning(self, chunk, variant, graph_chunk):
        self.chunk = int(chunk)
        self.variant = None if variant is None else int(variant)
        self.graph_chunk = None if graph_chunk is None else int(graph_chunk)

    @classmethod
    def from_device(cls, verts, neighbors, order="sfc", tiled=True, stream=None, steps=4,
                    chunk=0, variant=None, graph_chunk=None):
        """The same graph from device tensors (verts float64 [V,3], neighbours
        int32 [V,6]), built on the device (nx_sfc_order_device,
        nx_ell_from_adjacency_device); the tile plan is cut on the host from
        copies of the device tables."""
        import torch
        from . import util
        V = verts.shape[0]
        if neighbors.shape != (V, 6) or neighbors.dtype != torch.int32:
            raise ValueError("neighbors must be an int32 [V,6] device tensor")
        if isinstance(order, str):
            if order != "sfc":
                raise ValueError(f"unknown order {order!r}")
            order = util.sfc_order_device(verts, stream)
        self = cls.__new__(cls)
        self.V = V
        self.device = verts.device
        col = torch.empty((6, V), dtype=torch.int32, device=verts.device)
        elen = torch.empty((6, V), dtype=torch.float64, device=verts.device)
        ws = torch.empty(int(_lib.lib().nx_ell_workspace_bytes(V)), dtype=torch.uint8,
                         device=verts.device)
        ctx = _lib.Context.get(verts.device)
        _lib.check(_lib.lib().nx_ell_from_adjacency_device(
            ctx.handle, _lib.ptr(neighbors.contiguous()), V, _lib.ptr(verts.contiguous()),
            _lib.ptr(order), _lib.ptr(col), _lib.ptr(elen), _lib.ptr(ws),
            _lib.stream_ptr(stream)))
        ctx.sync_check(stream)
        self.col, self.elen = col, elen
        self._odev = order
        self.order = order.cpu().numpy() if order is not None else None
        self.plan = None
        self.stream_bytes_per_launch = None
        self._plan_blob = None
        self.steps = int(steps)
        self._tuning(chunk, variant, graph_chunk)
        if tiled:
            self._build_plan(col.cpu().numpy(), elen.cpu().numpy())
        return self

    def _build_plan(self, col, elen):
        import torch
        L = _lib.lib()
        plan = _lib.NxTilePlan()
        rc = L.nx_tile_plan_build_chunk(_lib.ptr(col), _lib.ptr(elen), self.V, self.steps,
                                        self.chunk, None, 0, ctypes.byref(plan))
        if rc == _lib.NX_ERANGE and self.steps != 2:
            self.steps = 2          # tiles too wide for four rings: two steps per launch
            rc = L.nx_tile_plan_build_chunk(_lib.ptr(col), _lib.ptr(elen), self.V, self.steps,
                                            self.chunk, None, 0, ctypes.byref(plan))
        if rc == _lib.NX_ERANGE:
            return
        _lib.check(rc)
        img = np.zeros(plan.bytes, dtype=np.uint8)
        _lib.check(L.nx_tile_plan_build_chunk(_lib.ptr(col), _lib.ptr(elen), self.V, self.steps,
                                              self.chunk, _lib.ptr(img), plan.bytes,
                                              ctypes.byref(plan)))
        self.stream_bytes_per_launch = _plan_stream_bytes(plan, img, self.V)
        blob = torch.from_numpy(img).to(self.device)
        if blob.data_ptr() % 256:
            raise RuntimeError("tile plan copy is not 256-byte aligned")
        plan.blob = blob.data_ptr()
        self.plan = plan
        self._plan_blob = blob

    @property
    def nbytes(self):
        n = self.col.numel() * self.col.element_size() + self.elen.numel() * self.elen.element_size()
        return n + (self._plan_blob.numel() if self._plan_blob is not None else 0)

    def _order_dev(self):
        import torch
        if getattr(self, "_odev", None) is None:
            self._odev = torch.from_numpy(self.order).to(self.device)
        return self._odev

    def permute_device(self, t):
        """Caller-numbered device tensor [V] -> device numbering (a copy)."""
        return t.clone() if self.order is None else t[self._order_dev()]

    def unpermute_device(self, t):
        """Device-numbered device tensor [V] -> caller numbering (a copy)."""
        import torch
        if self.order is None:
            return t.clone()
        out = torch.empty_like(t)
        out[self._order_dev()] = t
        return out

    def permute(self, a):
        """Caller-numbered per-vertex array -> device numbering."""
        return a if self.order is None else a[self.order]

    def unpermute(self, a):
        """Device-numbered per-vertex array -> caller numbering."""
        if self.order is None:
            return a
        out = np.empty_like(a)
        out[self.order] = a
        return out


class ErosionState:
    """Device-resident h/w/s (FP64, double-buffered) of one planet."""

    def __init__(self, V, device=None):
        import torch
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.V = int(V)
        nbytes = int(_lib.lib().nx_erosion_state_bytes(self.V))
        # 256-byte alignment of the workspace: torch's caching allocator
        # returns 512-byte aligned blocks.
        self.buf = torch.empty(nbytes, dtype=torch.uint8, device=device)
        if self.buf.data_ptr() % 256:
            raise RuntimeError("erosion workspace is not 256-byte aligned")
        self.device = self.buf.device
        self.ctx = _lib.Context.get(self.device)

    @property
    def nbytes(self):
        return self.buf.numel()

    def load(self, h, w=None, s=None, stream=None):
        """h, w, s: float64 [V] device tensors (w, s None for zeros)."""
        for t in (h, w, s):
            if t is not None and (t.dtype.itemsize != 8 or t.numel() != self.V
                                  or not t.is_contiguous()):
                raise ValueError("fields must be contiguous float64 [V] tensors")
        _lib.check(_lib.lib().nx_erosion_load(self.ctx.handle, _lib.ptr(self.buf), self.V,
                                              _lib.ptr(h), _lib.ptr(w), _lib.ptr(s),
                                              _lib.stream_ptr(stream)))

    def run(self, graph, iters, rain=RAIN_DEFAULT, stream=None, rule=3):
        """iters steps, rain added before each; rule 3 is erosion_iteration3
        (erosion.py:197-279), rule 4 erosion_iteration4 (erosion.py:301-381)."""
        if graph.V != self.V:
            raise ValueError("graph and state sizes differ")
        plan = ctypes.byref(graph.plan) if graph.plan is not None else None
        variant = getattr(graph, "variant", None)
        four = graph.plan is not None and graph.plan.steps == 4
        with self.ctx.options(graph_chunk=getattr(graph, "graph_chunk", None),
                              pair_kernel=None if four else variant,
                              multi_kernel=variant if four else None):
            _lib.check(_lib.lib().nx_erosion_run_rule(
                self.ctx.handle, _lib.ptr(self.buf), self.V, _lib.ptr(graph.col),
                _lib.ptr(graph.elen), plan, int(iters), float(rain), int(rule),
                _lib.stream_ptr(stream)))

    def store(self, h=None, w=None, s=None, stream=None):
        _lib.check(_lib.lib().nx_erosion_store(self.ctx.handle, _lib.ptr(self.buf), self.V,
                                               _lib.ptr(h), _lib.ptr(w), _lib.ptr(s),
                                               _lib.stream_ptr(stream)))

    def mass(self, stream=None):
        """(sum(h+s), sum(|h|+|s|), sum(w)) in FP64, as a device tensor [3]."""
        import torch
        out = torch.empty(3, dtype=torch.float64, device=self.device)
        _lib.check(_lib.lib().nx_erosion_mass(self.ctx.handle, _lib.ptr(self.buf), self.V,
                                              _lib.ptr(out), _lib.stream_ptr(stream)))
        return out


def _digest(a):
    """Content digest of a contiguous array (128-bit xxh3)."""
    import xxhash
    a = np.ascontiguousarray(a)
    return (a.shape, a.dtype.str, xxhash.xxh3_128_hexdigest(memoryview(a).cast("B")))


_GRAPHS = {}        # (digest of verts, digest of neighbours, options) -> ErosionGraph
_GRAPHS_MAX = 2


def graph_for(verts, neighbors, order="sfc", tiled=True):
    """The ErosionGraph of (verts, neighbors), built once and reused while the
    arrays hold the same contents.  The reference's loops call the step
    once per iteration with the same mesh (erosion.py:180-184), so the
    drop-ins keep the device tables, vertex order and tile plan of the last
    two meshes instead of rebuilding them per call; the key is the arrays'
    contents, so a mesh edited in place gets a new graph."""
    import torch
    dev = torch.cuda.current_device()
    key = (_digest(np.asarray(verts, dtype=np.float64)),
           _digest(np.asarray(neighbors, dtype=np.int32)), order, bool(tiled), dev)
    g = _GRAPHS.pop(key, None)
    if g is None:
        from . import autotune
        t = autotune.lookup_erosion(np.shape(neighbors)[0]) if tiled else None
        g = ErosionGraph(verts, neighbors, order=order, tiled=tiled,
                         **(t.graph_kwargs() if t is not None else {}))
        while len(_GRAPHS) >= _GRAPHS_MAX:
            _GRAPHS.pop(next(iter(_GRAPHS)))
    _GRAPHS[key] = g          # most recently used last
    return g


def _dev(a, graph):
    import torch
    return torch.from_numpy(np.ascontiguousarray(graph.permute(a), dtype=np.float64)).to(
        graph.device)


def _snapshot_image(h_dev, key):
    """erosion.py:186-192 on the device: the heights rescaled to 0..255,
    build_image_data's gray image of them (util.py:369-438: a float64 layer
    is rescaled to 0..255 once more, sampled per pixel by make_gray_array and
    cast to uint8) and save_image(..., cfg.SNAP_DIR, "erosion_snapshot").
    The reference puts the bare array into the dict where build_image_data
    expects [array, mode]; the layer is written as the 'gray' image that
    call is meant to produce."""
    import os
    import torch
    from PIL import Image
    from . import cfg, util
    if cfg.SNAP_DIR is None or cfg.IMG_QUERY_DATA is None:
        raise ValueError("erosion snapshots need cfg.SNAP_DIR and cfg.IMG_QUERY_DATA")
    dists, nbrs = cfg.IMG_QUERY_DATA
    if not isinstance(dists, torch.Tensor):
        dists = torch.from_numpy(np.ascontiguousarray(dists, dtype=np.float64)).to(h_dev.device)
        nbrs = torch.from_numpy(np.ascontiguousarray(nbrs, dtype=np.int64)).to(h_dev.device)
        cfg.IMG_QUERY_DATA = (dists, nbrs)   # keep the device copies for later passes
    if tuple(dists.shape[:2]) != (int(cfg.IMG_HEIGHT), int(cfg.IMG_WIDTH)):
        raise ValueError("cfg.IMG_QUERY_DATA does not match cfg.IMG_WIDTH x cfg.IMG_HEIGHT")
    layer, _ = util.rescale_device(h_dev, 0, 255)            # erosion.py:188
    layer, _ = util.rescale_device(layer, 0, 255, out=layer)  # util.py:397-398
    pixels = util.make_gray_array_device(dists, nbrs, layer).to(torch.uint8)
    _lib.Context.get(h_dev.device).sync_check()
    Image.fromarray(pixels.cpu().numpy()).save(
        os.path.join(cfg.SNAP_DIR, f"erosion_snapshot_{key}.{cfg.IMG_FORMAT}"))


def erode_terrain3(nodes, neighbors, heights, num_iter=1, snapshot=False, rain=RAIN_DEFAULT):
    """erosion.py:171-192: erode ``heights`` (float64[V]) in place.

    Water and sediment start at zero and are discarded, as in the reference.
    With ``snapshot``, each pass also writes the gray image of the heights
    rescaled to 0..255 to cfg.SNAP_DIR as erosion_snapshot_NNN (NNN the pass
    number from 001), sampled through cfg.IMG_QUERY_DATA.
    ``rain`` is the per-iteration rainfall (the reference hard-codes 0.3/320).
    """
    import torch
    if heights.dtype != np.float64:
        raise TypeError("heights must be float64")
    if num_iter <= 0:
        num_iter = 1
    graph = graph_for(nodes, neighbors)
    st = ErosionState(graph.V, graph.device)
    st.load(_dev(heights, graph))
    h = torch.empty(graph.V, dtype=torch.float64, device=graph.device)
    if snapshot:
        for i in range(num_iter):
            st.run(graph, 1, rain)
            st.store(h)
            _snapshot_image(graph.unpermute_device(h), f"{i + 1:03d}")
    else:
        st.run(graph, num_iter, rain)
        st.store(h)
    st.ctx.sync_check()
    heights[...] = graph.unpermute(h.cpu().numpy())


def erosion_iteration3(verts, neighbors, r_buff, wat, sed):
    """erosion.py:197-279: one step; r_buff, wat, sed (float64[V]) updated in place."""
    import torch
    for a in (r_buff, wat, sed):
        if a.dtype != np.float64:
            raise TypeError("fields must be float64")
    graph = graph_for(verts, neighbors)
    st = ErosionState(graph.V, graph.device)
    st.load(_dev(r_buff, graph), _dev(wat, graph), _dev(sed, graph))
    st.run(graph, 1, 0.0)
    outs = [torch.empty(graph.V, dtype=torch.float64, device=graph.device) for _ in range(3)]
    st.store(*outs)
    st.ctx.sync_check()
    for a, t in zip((r_buff, wat, sed), outs):
        a[...] = graph.unpermute(t.cpu().numpy())


def erosion_iteration4(verts, neighbors, r_buff, wat, sed):
    """erosion.py:301-381: one clamped step; r_buff, wat, sed (float64[V])
    updated in place."""
    import torch
    for a in (r_buff, wat, sed):
        if a.dtype != np.float64:
            raise TypeError("fields must be float64")
    graph = graph_for(verts, neighbors)
    st = ErosionState(graph.V, graph.device)
    st.load(_dev(r_buff, graph), _dev(wat, graph), _dev(sed, graph))
    st.run(graph, 1, 0.0, rule=4)
    outs = [torch.empty(graph.V, dtype=torch.float64, device=graph.device) for _ in range(3)]
    st.store(*outs)
    st.ctx.sync_check()
    for a, t in zip((r_buff, wat, sed), outs):
        a[...] = graph.unpermute(t.cpu().numpy())


def erode_terrain4(nodes, neighbors, heights, num_iter=1, snapshot=None):
    """erosion.py:282-296.  As in the reference, this driver adds 0.3/320 of
    rain and runs erosion_iteration3 (not erosion_iteration4) each pass, so
    it is erode_terrain3 without the snapshot option."""
    erode_terrain3(nodes, neighbors, heights, num_iter, False, 0.3 / 320)


def _v1_run(graph, h_dev, iters):
    import torch
    tmp = torch.empty_like(h_dev)
    ctx = _lib.Context.get(h_dev.device)
    _lib.check(_lib.lib().nx_erosion1_run(ctx.handle, _lib.ptr(h_dev), _lib.ptr(tmp), graph.V,
                                          _lib.ptr(graph.col), int(iters), _lib.stream_ptr()))
    ctx.sync_check()


def erosion_iteration1(neighbors, r_buff, w_buff):
    """erosion.py:76-99: w_buff (float64[V]) receives one step from r_buff
    and is returned.  The step needs only the neighbour table."""
    if w_buff.dtype != np.float64:
        raise TypeError("w_buff must be float64")
    nb = np.ascontiguousarray(neighbors, dtype=np.int32)
    graph = graph_for(np.zeros((nb.shape[0], 3)), nb, order=None, tiled=False)
    h = _dev(r_buff, graph)
    _v1_run(graph, h, 1)
    w_buff[...] = graph.unpermute(h.cpu().numpy())
    return w_buff


def erode_terrain1(nodes, neighbors, heights, num_iter=1, snapshot=None):
    """erosion.py:42-70: num_iter steps of erosion_iteration1 (at least
    one), heights (float64[V]) updated in place and returned, as the
    reference returns its read buffer."""
    if heights.dtype != np.float64:
        raise TypeError("heights must be float64")
    if num_iter <= 0:
        num_iter = 1
    graph = graph_for(nodes, neighbors, tiled=False)
    h = _dev(heights, graph)
    _v1_run(graph, h, num_iter)
    heights[...] = graph.unpermute(h.cpu().numpy())
    return heights
