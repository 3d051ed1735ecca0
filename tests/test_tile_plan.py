"""The tile plan for K erosion steps per launch (nx_tile_plan_build_steps,
host), K = 2 and 4.

Checked on the CPU: tiles cover the vertices in order, each tile's rings
are exactly its successive layers of outside neighbours, every local
neighbour index names the same vertex as the global table, edge lengths are
carried over, and K Jacobi steps computed from the plan's local view (step j
on the tile and rings 1..K-j) equal K steps of the FP64 oracle bit for bit.
A numbering without locality is refused.
"""
import ctypes

import numpy as np
import pytest

from nixis_amd import _lib, util

EVAP, SOL, CAP = 0.1 / 320, 0.01 / 320, 0.2 / 320


def _ell(d, order="sfc"):
    pts, nb = util.icosphere_with_neighbors(d)
    V = pts.shape[0]
    o = util.sfc_order(pts) if isinstance(order, str) else order
    col = np.empty((6, V), np.int32)
    elen = np.empty((6, V), np.float64)
    _lib.check(_lib.lib().nx_ell_from_adjacency(_lib.ptr(nb), V, _lib.ptr(pts), _lib.ptr(o),
                                                _lib.ptr(col), _lib.ptr(elen)))
    return pts, nb, o, col, elen


def _plan(col, elen, steps=2, chunk=0):
    V = col.shape[1]
    L = _lib.lib()
    q = _lib.NxTilePlan()
    rc = L.nx_tile_plan_build_chunk(_lib.ptr(col), _lib.ptr(elen), V, steps, chunk, None, 0,
                                    ctypes.byref(q))
    _lib.check(rc)
    img = np.zeros(q.bytes, np.uint8)
    _lib.check(L.nx_tile_plan_build_chunk(_lib.ptr(col), _lib.ptr(elen), V, steps, chunk,
                                          _lib.ptr(img), q.bytes, ctypes.byref(q)))
    assert q.steps == steps
    nt = q.ntiles

    def arr(off, dt, n):
        return np.frombuffer(img, dtype=dt, count=n, offset=off)

    start = arr(q.off_start, np.int32, nt + 1)
    cs = arr(q.off_cs, np.int32, nt)
    load = arr(q.off_load, np.int32, nt)
    rbase = arr(q.off_rbase, np.int64, nt + 1)
    ring = arr(q.off_ring, np.int32, int(rbase[-1]))
    cbase = arr(q.off_cbase, np.int64, nt + 1)
    lcol = arr(q.off_lcol, np.uint16, 6 * int(cbase[-1]))
    ebase = arr(q.off_ebase, np.int64, nt + 1)
    lel = arr(q.off_lel, np.float64, int(ebase[-1]))
    pref = arr(q.off_pref, np.int32, (steps + 1) * nt).reshape(nt, steps + 1)
    return q, start, cs, load, rbase, ring, cbase, lcol, (lel, ebase), pref


FULL = 3   # length columns that hold every step-1 vertex (the tile's own hold no more)


def _where(colm, owner, c, T):
    """Index in a tile's length block of column colm of owner: columns below
    FULL over all c step-1 vertices, later ones over the ring vertices
    T..c-1 only."""
    return np.where(colm < FULL, colm * c + owner, FULL * c + (colm - FULL) * (c - T) + owner - T)


def _decode(codes, lens, c, T):
    """Slot codes [6][c] of one tile -> (local neighbour [6][c], the length
    each slot reads [6][c]): bits 0-11 neighbour, 12-14 column, bit 15 set
    when the slot's own vertex holds the length, else the neighbour does."""
    codes = codes.astype(np.int64)
    nbr = codes & 4095
    colm = (codes >> 12) & 7
    mine = (codes >> 15) & 1
    li = np.broadcast_to(np.arange(c), codes.shape)
    owner = np.where(mine == 1, li, nbr)
    assert (owner < c).all() and (owner[colm >= FULL] >= T).all()
    k = int(colm.max()) + 1
    assert lens.size == min(k, FULL) * c + max(k - FULL, 0) * (c - T)
    return nbr, lens[_where(colm, owner, c, T)]


def _tiles(col, elen, steps=2, chunk=0):
    q, start, cs, load, rbase, ring, cbase, lcol, (lel, ebase), pref = _plan(col, elen, steps,
                                                                             chunk)
    for t in range(q.ntiles):
        a, b = int(start[t]), int(start[t + 1])
        T = b - a
        rg = ring[rbase[t]:rbase[t + 1]]
        glob = np.concatenate([np.arange(a, b), rg])          # local -> global
        c = int(cs[t])
        lens = lel[ebase[t]:ebase[t + 1]]
        lc, le = _decode(lcol[6 * cbase[t]:6 * cbase[t + 1]].reshape(6, c), lens, c, T)
        assert pref[t, 0] == T and pref[t, -2] == c and pref[t, -1] == int(load[t])
        yield q, a, b, T, c, int(load[t]), glob, lc, le, pref[t]


@pytest.mark.parametrize("d", [2, 8, 40])
def test_plan_covers_and_rings_are_exact(d):
    pts, nb, o, col, elen = _ell(d)
    V = col.shape[1]
    nxt = 0
    step1 = 0
    for q, a, b, T, c, L, glob, lc, le, _ in _tiles(col, elen):
        # two-step tiles: at most 768 owned, tile + ring 1 within one
        # 1024-thread block, tile + both rings within 2048
        assert a == nxt and 0 < T <= 768 and c <= 1024 and L <= 2048 and q.max_tile <= 768
        step1 = max(step1, c)
        nxt = b
        inside = set(range(a, b))
        r1 = sorted(set(col[:, a:b].ravel()) - inside)
        r2 = sorted(set(col[:, r1].ravel()) - inside - set(r1)) if r1 else []
        assert list(glob[T:c]) == r1 and list(glob[c:L]) == r2
        # every local neighbour of tile + ring 1 names the global neighbour,
        # and every slot but a pad (a pentagon's sixth slot, the vertex
        # itself: a tie, whatever the length) reads its edge's length
        assert (glob[lc] == col[:, glob[:c]]).all()
        pad = lc == np.arange(c)[None, :]
        assert (le == elen[:, glob[:c]])[~pad].all()
    assert nxt == V and q.max_step1 == step1


@pytest.mark.parametrize("d,steps", [(40, 2), (100, 2), (40, 4)])
def test_edge_lengths_are_stored_once(d, steps):
    """Each undirected edge between two step-1 vertices of a tile has its
    length in one place, read by both ends; the tile's own vertices hold at
    most FULL lengths each, so only the ring vertices reach the later,
    shorter columns, and the block holds few padding entries."""
    pts, nb, o, col, elen = _ell(d)
    q, start, cs, load, rbase, ring, cbase, lcol, (lel, ebase), pref = _plan(col, elen, steps)
    ncols, stored, owned = [], 0, 0
    for t in range(q.ntiles):
        c = int(cs[t])
        T = int(start[t + 1]) - int(start[t])
        codes = lcol[6 * cbase[t]:6 * cbase[t + 1]].reshape(6, c).astype(np.int64)
        nbr, colm, mine = codes & 4095, (codes >> 12) & 7, (codes >> 15) & 1
        k = int(colm.max()) + 1
        ncols.append(k)
        blk = int(ebase[t + 1]) - int(ebase[t])
        assert blk == min(k, FULL) * c + max(k - FULL, 0) * (c - T)
        li = np.broadcast_to(np.arange(c), codes.shape)
        owner = np.where(mine == 1, li, nbr)
        assert (colm[:, :T][mine[:, :T] == 1] < FULL).all()     # own vertices: at most FULL
        where = _where(colm, owner, c, T)
        stored += blk
        owned += int(((mine == 1) & (nbr != li)).sum())
        for j in range(c):
            for s in range(6):
                n = int(nbr[s, j])
                if n == j or n >= c:
                    assert mine[s, j] == 1      # pads and edges leaving the set
                    continue
                back = [r for r in range(6) if nbr[r, n] == j]
                assert len(back) == 1 and where[s, j] == where[back[0], n]
        # a column entry is read by at most two slots (the edge's two ends)
        used = np.bincount(where[(nbr != li)].ravel(), minlength=blk)
        assert used.max() <= 2 and used.size == blk
    ncols = np.array(ncols)
    print(f"\nd={d} steps={steps}: {q.ntiles} tiles, columns {np.bincount(ncols)}, "
          f"padding {1 - owned / stored:.1%}")
    assert (ncols <= 6).all() and 1 - owned / stored <= 0.15


def _step_rows(n_rows, Hs, Ws, Ss, lc, d):
    """One step of the first n_rows local vertices (erosion.py:209-267)."""
    Hn, Wn, Sn = Hs[:n_rows].copy(), Ws[:n_rows].copy(), Ss[:n_rows].copy()
    for i in range(n_rows):
        S, W = Ss[i], Ws[i]
        for k in range(6):
            n = lc[k, i]
            if Hs[n] > Hs[i]:
                S += SOL * Ws[n]
                W += Ws[n] * d[k, i]
            elif Hs[n] < Hs[i]:
                S -= SOL * Ws[n]
                W -= Ws[n] * d[k, i]
        Hn[i] = Hs[i] - S
        Sn[i] = Ss[i] + S
        Wn[i] = Ws[i] + (W - W * EVAP)
        if Sn[i] > CAP * Wn[i]:
            Hn[i] += Sn[i] - CAP * Wn[i]
            Sn[i] -= Sn[i] - CAP * Wn[i]
    return Hn, Wn, Sn


def _two_steps_local(h, w, s, glob, c, T, lc, d):
    """Two steps of the tile from its local view: step 1 on tile + ring 1
    from the staged fields of tile + both rings, step 2 on the tile, with
    the plan's local neighbour indices and edge lengths."""
    H, Wt, Sd = h[glob].copy(), w[glob].copy(), s[glob].copy()

    def step(n_rows, Hs, Ws, Ss):
        Hn, Wn, Sn = Hs[:n_rows].copy(), Ws[:n_rows].copy(), Ss[:n_rows].copy()
        for i in range(n_rows):
            S, W = Ss[i], Ws[i]
            for k in range(6):
                n = lc[k, i]
                if Hs[n] > Hs[i]:
                    S += SOL * Ws[n]
                    W += Ws[n] * d[k, i]
                elif Hs[n] < Hs[i]:
                    S -= SOL * Ws[n]
                    W -= Ws[n] * d[k, i]
            Hn[i] = Hs[i] - S
            Sn[i] = Ss[i] + S
            Wn[i] = Ws[i] + (W - W * EVAP)
            if Sn[i] > CAP * Wn[i]:
                Hn[i] += Sn[i] - CAP * Wn[i]
                Sn[i] -= Sn[i] - CAP * Wn[i]
        return Hn, Wn, Sn

    H1, W1, S1 = step(c, H, Wt, Sd)
    H2, W2, S2 = step(T, H1, W1, S1)
    return H2, W2, S2


def test_two_steps_from_the_plan_equal_the_oracle(oracle):
    d = 6
    pts, nb, o, col, elen = _ell(d)
    V = col.shape[1]
    P = pts[o]
    # oracle neighbour table in device order, -1 pads where col points at self
    nbd = col.T.copy()
    nbd[nbd == np.arange(V)[:, None]] = -1
    rng = np.random.default_rng(2)
    h = rng.uniform(-9000, 9000, V)
    w = rng.uniform(0, 1, V)
    s = rng.uniform(0, 0.01, V)
    hr, wr, sr = h.copy(), w.copy(), s.copy()
    for _ in range(2):
        oracle.erosion_iteration3(P, nbd, hr, wr, sr)
    for q, a, b, T, c, L, glob, lc, le, _ in _tiles(col, elen):
        H2, W2, S2 = _two_steps_local(h, w, s, glob, c, T, lc, le)
        assert (H2 == hr[a:b]).all() and (W2 == wr[a:b]).all() and (S2 == sr[a:b]).all()


@pytest.mark.parametrize("d,chunk", [(3, 0), (12, 0), (40, 0), (40, 1536), (60, 2048)])
def test_four_step_rings_are_exact_layers(d, chunk):
    """Four-step tiles from the default chunk and from chunks above the
    two-step limit of 1024: the rings are the tile's successive layers of
    outside neighbours, tile + rings 1..3 stay within 2048 vertices and tile
    + rings within 2688."""
    pts, nb, o, col, elen = _ell(d)
    V = col.shape[1]
    nxt = 0
    for q, a, b, T, c, L, glob, lc, le, pref in _tiles(col, elen, 4, chunk):
        assert a == nxt and c <= 2048 and L <= 2688 and T <= max(chunk, 1152)
        nxt = b
        seen = set(range(a, b))
        layer = list(range(a, b))
        for m in range(4):
            nxt_layer = sorted(set(col[:, layer].ravel()) - seen) if layer else []
            assert list(glob[pref[m]:pref[m + 1]]) == nxt_layer
            seen |= set(nxt_layer)
            layer = nxt_layer
        pad = lc == np.arange(c)[None, :]
        assert (glob[lc] == col[:, glob[:c]]).all() and (le == elen[:, glob[:c]])[~pad].all()
    assert nxt == V


def test_four_steps_from_the_plan_equal_the_oracle(oracle):
    d = 10
    pts, nb, o, col, elen = _ell(d)
    V = col.shape[1]
    P = pts[o]
    nbd = col.T.copy()
    nbd[nbd == np.arange(V)[:, None]] = -1
    rng = np.random.default_rng(4)
    h = rng.uniform(-9000, 9000, V)
    w = rng.uniform(0, 1, V)
    s = rng.uniform(0, 0.01, V)
    hr, wr, sr = h.copy(), w.copy(), s.copy()
    for _ in range(4):
        oracle.erosion_iteration3(P, nbd, hr, wr, sr)
    for q, a, b, T, c, L, glob, lc, le, pref in _tiles(col, elen, 4):
        H, W, S = h[glob].copy(), w[glob].copy(), s[glob].copy()
        for j in range(1, 5):
            n = int(pref[4 - j])
            Hn, Wn, Sn = _step_rows(n, H, W, S, lc, le)
            H[:n], W[:n], S[:n] = Hn, Wn, Sn
        assert (H[:T] == hr[a:b]).all() and (W[:T] == wr[a:b]).all() and (S[:T] == sr[a:b]).all()


@pytest.mark.parametrize("n_steps", [2, 3])
def test_tail_steps_from_a_four_step_plan_equal_the_oracle(oracle, n_steps):
    """A run's last two or three steps on a four-step plan: step j covers
    the first pref[n - j] local vertices (tile + rings 1..n-j), reading only
    the first pref[n] (the vertices a tail launch stages; the rest are set
    to NaN here), and equals n steps of the oracle bit for bit."""
    d = 16
    pts, nb, o, col, elen = _ell(d)
    V = col.shape[1]
    P = pts[o]
    nbd = col.T.copy()
    nbd[nbd == np.arange(V)[:, None]] = -1
    rng = np.random.default_rng(n_steps)
    h = rng.uniform(-9000, 9000, V)
    w = rng.uniform(0, 1, V)
    s = rng.uniform(0, 0.01, V)
    hr, wr, sr = h.copy(), w.copy(), s.copy()
    for _ in range(n_steps):
        oracle.erosion_iteration3(P, nbd, hr, wr, sr)
    for q, a, b, T, c, L, glob, lc, le, pref in _tiles(col, elen, 4, 1024):
        H, W, S = h[glob].copy(), w[glob].copy(), s[glob].copy()
        staged = int(pref[n_steps])
        H[staged:], W[staged:], S[staged:] = np.nan, np.nan, np.nan
        for j in range(1, n_steps + 1):
            n = int(pref[n_steps - j])
            Hn, Wn, Sn = _step_rows(n, H, W, S, lc, le)
            H[:n], W[:n], S[:n] = Hn, Wn, Sn
        assert (H[:T] == hr[a:b]).all() and (W[:T] == wr[a:b]).all() and (S[:T] == sr[a:b]).all()


def test_chunk_limits_by_step_count():
    """Two-step plans start from at most 1024 vertices (the pair kernel's
    tile limit), four-step plans from at most 2048."""
    _, _, _, col, elen = _ell(12)
    V = col.shape[1]
    q = _lib.NxTilePlan()
    L = _lib.lib()
    for steps, chunk, ok in ((2, 1024, True), (2, 1025, False), (4, 2048, True),
                             (4, 2049, False), (4, 31, False)):
        rc = L.nx_tile_plan_build_chunk(_lib.ptr(col), _lib.ptr(elen), V, steps, chunk, None, 0,
                                        ctypes.byref(q))
        assert (rc == _lib.NX_OK) == ok, (steps, chunk, rc)
        if not ok:
            assert rc == _lib.NX_EINVAL


def test_numbering_without_locality_is_refused():
    pts, nb, _, _, _ = _ell(40, order=None)
    V = pts.shape[0]
    o = np.random.default_rng(0).permutation(V).astype(np.int64)
    _, _, _, col, elen = _ell(40, order=o)
    q = _lib.NxTilePlan()
    rc = _lib.lib().nx_tile_plan_build(_lib.ptr(col), _lib.ptr(elen), V, None, 0, ctypes.byref(q))
    assert rc == _lib.NX_ERANGE
