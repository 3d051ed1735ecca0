"""GPU parity of the height post-processing after noise against the oracle
(util.py:110-254, util.py:556-566, terrain.py:61-72, nixis.py:330-364).

rescale (every form) and the elevation mask are FP64 maps in the
reference's operation order and must equal the oracle bit for bit.
power_rescale's bounds (the reference's sequential if/elif loop, done on the
device with a min-scan) must be exact; its values pass through `pow`.  The
device computes x**2 and x**0.5 as the correctly rounded t*t and sqrt(t);
glibc's pow, which the reference and the oracle call, misses the correctly
rounded value by 1 ulp for about 0.08% of inputs (and other powers use the
device's own pow), so those values are held to |gpu - oracle| <= 8 eps * (|oracle| + max|x|) per value
(a few ulp of t**p, scaled back by the mask range, which max|x| bounds).  The chain inherits that
bound on the final heights (8 eps * (max_alt - min_alt) per power step,
times the final rescale's gain); its mask and ocean level are exact.
"""
import numpy as np
import pytest
import torch

from nixis_amd import _lib, opensimplex as osx, terrain, util

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def _dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def _noise_heights(d=100, seed=7):
    pts = util.create_points(d)
    perm, pgi = osx.init(seed)
    return terrain.sample_octaves(pts, None, perm, pgi, 8, 1.5, 0.4, 2.0, 0.5)


@pytest.mark.parametrize("kw", [
    {},
    {"mid": 0.1},
    {"mid": -0.2, "mode": "lower"},
    {"mid": 0.3, "mode": "upper"},
    {"u_min": -5.0},
    {"u_max": 7.5, "u_min": -0.1},
    {"mid": 0.0, "u_min": -3.0, "u_max": 3.0},
])
def test_rescale_forms_bit_exact(oracle, kw):
    rng = np.random.default_rng(11)
    x = rng.standard_normal(200_003)
    print("rescale", kw, flush=True)
    got = util.rescale(x, -4000.0, 8850.0, **kw)
    ref = oracle.rescale_ex(x, -4000.0, 8850.0, **kw)
    assert got.dtype == np.float64 and np.array_equal(got, ref)


@pytest.mark.parametrize("special", ["nan", "-nan", "inf", "-inf"])
def test_rescale_special_values_match_oracle(oracle, special):
    """One NaN of either sign makes both extremes NaN, as numba's np.min and
    np.max do (the device keys order a NaN at one end only); an infinity is
    an extreme like any value.  Every form equals the oracle: NaN where the
    oracle has NaN, bit for bit elsewhere."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal(100_003)
    x[4321] = {"nan": np.nan, "-nan": -np.nan, "inf": np.inf, "-inf": -np.inf}[special]
    for kw in ({}, {"mid": 0.1}, {"mid": -0.2, "mode": "lower"}, {"mid": 0.3, "mode": "upper"},
               {"u_min": -5.0, "u_max": 5.0}):
        print(special, kw, flush=True)
        got = util.rescale(x, -4000.0, 8850.0, **kw)
        ref = oracle.rescale_ex(x, -4000.0, 8850.0, **kw)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        ok = ~np.isnan(ref)
        assert np.array_equal(got[ok], ref[ok])


def test_rescale_hand_worked_on_device():
    x = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    assert util.rescale(x, 0, 10, mid=2.0).tolist() == [0, 1, 2, 6, 10]
    assert util.rescale(x, -6, 0, mid=2.0, mode="lower").tolist() == [-6, -2, 2, 3, 4]
    assert util.rescale(x, 0, 10, mid=2.0, mode="upper").tolist() == [0, 1, 2, 6, 10]
    assert util.rescale(np.array([1.0, 2.0, 3.0]), 0, 4, u_max=5.0).tolist() == [0, 1, 2]
    assert util.rescale(x, 0, 10, mode="upper") is x          # mode without mid: x itself


def test_rescale_zero_ranges_raise():
    with pytest.raises(ZeroDivisionError):
        util.rescale(np.ones(1000), 0, 1)
    with pytest.raises(ZeroDivisionError):
        util.rescale(np.array([1.0, 2.0]), 0, 3, mid=1.0)    # mid - x_min == 0
    # a one-sided mode never divides by the other half's zero range
    out = util.rescale(np.array([1.0, 2.0]), 0, 3, mid=2.0, mode="lower")
    assert out.tolist() == [0.0, 2.0]


def test_mask_bit_exact(oracle):
    h = _noise_heights(60)
    for level in (float(np.median(h)), float(h.min()), float(h.max()), float(h[17])):
        got = terrain.make_bool_elevation_mask(h, level)
        assert got.dtype == np.bool_
        assert np.array_equal(got, oracle.make_bool_elevation_mask(h, level))


def _check_power(x, mask, mode, power, oracle):
    got = util.power_rescale(x, mask, mode, power)
    ref = oracle.power_rescale(x, mask, mode, power)
    sel = (mask != 0) if mode == 1 else (mask == 0)
    assert np.array_equal(got[~sel], ref[~sel])            # untouched values copied
    g, r = got[sel], ref[sel]
    assert np.array_equal(np.isnan(g), np.isnan(r))         # pow of a negative base
    ok = ~np.isnan(r)
    g, r = g[ok], r[ok]
    tol = 8 * EPS * (np.abs(r) + np.abs(x).max())
    excess = (np.abs(g - r) - tol).max() if r.size else -1.0
    err = np.abs(g - r).max() if r.size else 0.0
    print(f"mode {mode} power {power}: max |gpu - oracle| {err:.3e}, "
          f"exact {np.mean(g == r) if r.size else 1.0:.4f}", flush=True)
    assert excess <= 0
    return got, ref


def test_power_rescale_matches_oracle(oracle):
    h = util.rescale(_noise_heights(100), -4000.0, 8850.0)
    level = oracle.find_percent_val(h.min(), h.max(), 55.0)
    mask = (h <= level).astype(np.uint8)
    for mode, power in ((1, 0.5), (0, 2.0), (1, 3.0), (0, 0.25)):
        _check_power(h, mask, mode, power, oracle)


def test_power_rescale_sequential_bounds(oracle):
    """Orderings where the sequential loop's bounds differ from the plain
    min/max of the selected values: descending runs (every value a new
    running minimum), ties, and the hand-worked quirk case."""
    x = np.array([7.0, 5.0, 10.0, 0.0])
    m = np.array([1, 1, 0, 0], np.uint8)
    got = util.power_rescale(x, m, 1, 2.0)
    assert np.allclose(got, [0.16 * -5 + 5, 5.0, 10.0, 0.0], rtol=0, atol=1e-12)
    # breaks placed on a chunk boundary (2048 values) and on a per-thread run
    # boundary (8 values), everything selected
    for at in (3 * 2048, 3 * 2048 + 8, 5 * 2048 + 3):
        x = np.linspace(100.0, -100.0, 20_000)
        x[at] = x[at - 1]
        x[at + 500] = 150.0        # above max before it: the upper bound
        mask = np.ones(x.size, np.uint8)
        print("break at", at, flush=True)
        _check_power(x, mask, 1, 2.0, oracle)
    rng = np.random.default_rng(5)
    n = 300_001
    for case in range(4):
        if case == 0:
            x = np.linspace(10.0, -10.0, n)              # strictly descending
        elif case == 1:
            x = np.round(rng.standard_normal(n) * 4) / 4  # many ties
        elif case == 2:
            x = np.concatenate([np.linspace(5, -5, n // 2), rng.standard_normal(n - n // 2)])
        else:
            x = np.cumsum(rng.standard_normal(n))         # a random walk
        mask = (rng.random(n) < 0.6).astype(np.uint8)
        for mode, power in ((1, 0.5), (0, 2.0)):
            print("case", case, "mode", mode, flush=True)
            try:
                ref = oracle.power_rescale(x, mask, mode, power)
            except ZeroDivisionError:
                with pytest.raises(ZeroDivisionError):
                    util.power_rescale(x, mask, mode, power)
                continue
            _check_power(x, mask, mode, power, oracle)


def test_power_rescale_zero_range_and_no_selection():
    x = np.full(5000, 3.0)
    with pytest.raises(ZeroDivisionError):
        util.power_rescale(x, np.ones(5000, np.uint8), 1, 0.5)
    # nothing selected: nothing divides, a copy comes back
    out = util.power_rescale(x, np.ones(5000, np.uint8), 0, 0.5)
    assert np.array_equal(out, x)
    assert np.array_equal(util.power_rescale(x, None, None, 2.0), x)


def test_height_chain_matches_oracle(oracle):
    h0 = _noise_heights(100)
    ref_h, ref_mask, ref_level = oracle.height_chain(h0, -4000.0, 8850.0, 55.0)
    t = _dev(h0)
    out, mask, level = util.height_chain_device(t, -4000.0, 8850.0, 55.0)
    _lib.Context.get(t.device).sync_check()
    got = out.cpu().numpy()
    assert level.item() == ref_level
    assert np.array_equal(mask.cpu().numpy().astype(bool), ref_mask)
    tol = 64 * EPS * (8850.0 + 4000.0)
    err = np.abs(got - ref_h).max()
    print(f"chain: max |gpu - oracle| {err:.3e} m (bound {tol:.3e}), "
          f"exact {np.mean(got == ref_h):.4f}", flush=True)
    assert err <= tol
    assert got.min() == -4000.0 and got.max() == 8850.0


# byte offset of the two-pass form's flag in the chain workspace (PrHead::fast)
_FAST_AT = 224


def _chain_both_forms(x, percent=55.0, offset=0):
    """Run the chain in the two-pass and the five-pass form on copies of x
    placed `offset` elements into their buffers; returns per form the heights,
    mask, level, raised error (or None) and the two-pass flag."""
    out = []
    ctx = _lib.Context.get(0)
    try:
        for passes in (2, 5):
            ctx.set_chain_passes(passes)
            buf = torch.zeros(x.size + offset, dtype=torch.float64, device="cuda")
            h = buf[offset:]
            h.copy_(_dev(x))
            need = _lib.lib().nx_height_chain_workspace_bytes(x.size)
            ws = torch.zeros(need, dtype=torch.uint8, device="cuda")
            hh, m, lv = util.height_chain_device(h, -4000.0, 8850.0, percent, workspace=ws)
            err = None
            try:
                ctx.sync_check()
            except ZeroDivisionError as e:
                err = type(e)
            fast = int(ws[_FAST_AT:_FAST_AT + 4].cpu().numpy().view(np.int32)[0])
            out.append((hh.cpu().numpy(), m.cpu().numpy(), lv.item(), err, fast))
    finally:
        ctx.set_chain_passes(2)
    return out


def _same_bits(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("case", ["noise", "normal_odd_unaligned", "quantized", "few_values",
                                  "two_values"])
def test_height_chain_two_pass_equals_five_pass(case):
    """The two-pass form (a summary pass, the bounds derived from the
    summaries, one writing pass) gives the five-pass form's heights, mask,
    level and errors bit for bit; it runs whenever its conditions hold, and
    leaves degenerate inputs (a zero power_rescale range) to the five-pass
    kernels."""
    rng = np.random.default_rng(5)
    offset, want_fast = 0, 1
    if case == "noise":
        x = _noise_heights(100)
    elif case == "normal_odd_unaligned":
        x, offset = rng.standard_normal(1_000_001), 1
    elif case == "quantized":           # many equal values, some at the level
        x = rng.integers(0, 101, 300_000).astype(np.float64)
    elif case == "few_values":
        x = rng.integers(0, 4, 100_000).astype(np.float64)
    else:                               # the ocean is one value: range 0
        x, want_fast = np.tile([0.0, 1.0], 50_000), 0
    print("chain forms", case, x.size, offset, flush=True)
    (h2, m2, l2, e2, f2), (h5, m5, l5, e5, _) = _chain_both_forms(x, offset=offset)
    assert f2 == want_fast
    assert _same_bits(h2, h5) and np.array_equal(m2, m5) and l2 == l5 and e2 is e5
    if want_fast:
        assert e2 is None and h2.min() == -4000.0 and h2.max() == 8850.0
    else:
        assert e2 is ZeroDivisionError


@pytest.mark.parametrize("n", [1, 2, 3, 7, 511, 512, 513, 1025, 4097])
def test_small_and_ragged_sizes(oracle, n):
    """The min/max pass (per-wave runs of 512-value chunks), rescale and both
    forms of the height chain at sizes around the chunk and block edges:
    rescale bit for bit against the oracle, the chain's two forms bit for bit
    against each other, errors included (one value, or one per half, is a
    zero range)."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    print("small sizes", n, flush=True)
    if n > 1:
        got = util.rescale(x, -4000.0, 8850.0)
        assert np.array_equal(got, oracle.rescale_ex(x, -4000.0, 8850.0))
    (h2, m2, l2, e2, _), (h5, m5, l5, e5, _) = _chain_both_forms(x)
    assert _same_bits(h2, h5) and np.array_equal(m2, m5) and e2 is e5
    assert _same_bits(np.array([l2]), np.array([l5]))      # NaN when x is constant
    if e2 is None:
        ref_h, ref_mask, ref_level = oracle.height_chain(x, -4000.0, 8850.0, 55.0)
        assert l2 == ref_level and np.array_equal(m2.astype(bool), ref_mask)
        assert np.abs(h2 - ref_h).max() <= 64 * EPS * (8850.0 + 4000.0)
    else:
        with pytest.raises(ZeroDivisionError):
            oracle.height_chain(x, -4000.0, 8850.0, 55.0)


def test_height_chain_full_size_properties():
    """At 10,000,002 vertices (d=1000): the final range is exactly
    [min_alt, max_alt], the ocean is exactly the vertices at or below 0,
    and the ocean level is the 55% point of the first rescale's range."""
    pts = torch.from_numpy(util.create_points(1000)).cuda()
    perm, pgi = osx.init(3)
    r, s = terrain.octave_scalars(8, 1.5, 0.4, 2.0, 0.5)
    h = torch.zeros(pts.shape[0], dtype=torch.float64, device="cuda")
    terrain.sample_octaves_device(pts, h, perm, pgi, r, s, 1.0)
    out, mask, level = util.height_chain_device(h, -4000.0, 8850.0, 55.0)
    _lib.Context.get(h.device).sync_check()
    assert out.min().item() == -4000.0 and out.max().item() == 8850.0
    assert level.item() == util.find_percent_val(-4000.0, 8850.0, 55.0)
    m = mask.bool()
    assert bool((out[m] <= 0).all()) and bool((out[~m] > 0).all())
    assert 0.3 < m.float().mean().item() < 0.8


def test_unaligned_arrays_take_the_scalar_path(oracle):
    """Views one element in (doubles 8- not 16-byte aligned, the mask at an
    odd address) give the same results as aligned arrays."""
    h0 = util.rescale(_noise_heights(40), -4000.0, 8850.0)
    n = h0.size
    buf = torch.zeros(n + 1, dtype=torch.float64, device="cuda")
    x = buf[1:]
    x.copy_(_dev(h0))
    assert x.data_ptr() % 16 == 8
    mbuf = torch.zeros(n + 1, dtype=torch.uint8, device="cuda")
    m = mbuf[1:]
    level = oracle.find_percent_val(h0.min(), h0.max(), 55.0)
    terrain.make_bool_elevation_mask_device(x, level, out=m)
    ref_m = oracle.make_bool_elevation_mask(h0, level)
    assert np.array_equal(m.cpu().numpy().astype(bool), ref_m)
    out = torch.empty_like(buf)[1:]
    util.power_rescale_device(x, m, 1, 0.5, out=out)
    ref = oracle.power_rescale(h0, ref_m.astype(np.uint8), 1, 0.5)
    got = out.cpu().numpy()
    assert np.abs(got - ref).max() <= 8 * EPS * np.abs(h0).max() * 2
    r, _ = util.rescale_device(x, -1.0, 1.0, out=out, mid=0.0)
    _lib.Context.get(x.device).sync_check()
    assert np.array_equal(r.cpu().numpy(), oracle.rescale_ex(h0, -1.0, 1.0, mid=0.0))
    raw = _noise_heights(40)
    x.copy_(_dev(raw))
    ref_h, ref_mask, ref_level = oracle.height_chain(raw, -4000.0, 8850.0, 55.0)
    hh, mm, lv = util.height_chain_device(x, -4000.0, 8850.0, 55.0)
    _lib.Context.get(x.device).sync_check()
    assert lv.item() == ref_level
    assert np.array_equal(mm.cpu().numpy().astype(bool), ref_mask)
    assert np.abs(hh.cpu().numpy() - ref_h).max() <= 64 * EPS * (8850.0 + 4000.0)


def test_terrain_heights_fused_equals_two_calls(oracle):
    """Noise with the min/max reduced in its own kernel, then the chain,
    equals noise then the chain as two calls, bit for bit; and the oracle's
    chain on the same noise heights within the stated bound."""
    pts = util.create_points(120)
    perm, pgi = osx.init(21)
    r, s = terrain.octave_scalars(10, 1.5, 0.4, 2.0, 0.5)
    dv = _dev(pts)
    h, m, lv = terrain.terrain_heights_device(dv, perm, pgi, r, s, 1.0, -4000.0, 8850.0, 55.0)
    h2 = terrain.sample_octaves_device(dv, None, perm, pgi, r, s, 1.0)
    raw = h2.cpu().numpy()
    h2, m2, lv2 = util.height_chain_device(h2, -4000.0, 8850.0, 55.0)
    _lib.Context.get(dv.device).sync_check()
    assert torch.equal(h, h2) and torch.equal(m, m2) and lv.item() == lv2.item()
    ref_h, ref_mask, ref_level = oracle.height_chain(raw, -4000.0, 8850.0, 55.0)
    assert lv.item() == ref_level
    assert np.array_equal(m.cpu().numpy().astype(bool), ref_mask)
    assert np.abs(h.cpu().numpy() - ref_h).max() <= 64 * EPS * (8850.0 + 4000.0)
