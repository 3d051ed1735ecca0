"""Pin the FP64 oracle's post-processing after noise on the CPU
(util.py:110-254, util.py:556-566, terrain.py:61-72, nixis.py:330-364).

Values are worked out by hand on small arrays: endpoints and midpoints of
each rescale form, the one-sided modes leaving the other side untouched,
u_min/u_max only widening the range, the mask's inclusive comparison, and
power_rescale including its sequential if/elif quirk (a value that is a new
running minimum is never considered for the maximum)."""
import numpy as np
import pytest


def test_rescale_standard_and_split(oracle):
    x = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    assert (oracle.rescale_ex(x, -10, 10) == [-10, -5, 0, 5, 10]).all()
    # mid=2: [0, 2] -> [0, 2] and (2, 4] -> (2, 10]
    assert (oracle.rescale_ex(x, 0, 10, mid=2.0) == [0, 1, 2, 6, 10]).all()
    # mid=1: lower half [0, 1] -> [-4, 1], upper half (1, 4] -> (1, 7]
    assert (oracle.rescale_ex(x, -4, 7, mid=1.0) == [-4, 1, 3, 5, 7]).all()


def test_rescale_one_sided_modes(oracle):
    x = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    lo = oracle.rescale_ex(x, -6, 0, mid=2.0, mode="lower")
    assert (lo == [-6, -2, 2, 3, 4]).all()          # <= mid: [0,2] -> [-6, 2]
    up = oracle.rescale_ex(x, 0, 10, mid=2.0, mode="upper")
    assert (up == [0, 1, 2, 6, 10]).all()           # >= mid: [2,4] -> [2, 10]
    same = oracle.rescale_ex(x, 0, 10, mode="upper")  # mode without mid: x itself
    assert same is not None and (same == x).all()


def test_rescale_user_bounds_only_widen(oracle):
    x = np.array([1.0, 2.0, 3.0])
    assert (oracle.rescale_ex(x, 0, 4, u_min=-1.0) == [2, 3, 4]).all()    # x_min -> -1
    assert (oracle.rescale_ex(x, 0, 4, u_min=2.0) == [0, 2, 4]).all()     # ignored
    assert (oracle.rescale_ex(x, 0, 4, u_max=5.0) == [0, 1, 2]).all()     # x_max -> 5


def test_rescale_zero_range_raises(oracle):
    with pytest.raises(ZeroDivisionError):
        oracle.rescale_ex(np.ones(4), 0, 1)
    with pytest.raises(ZeroDivisionError):
        oracle.rescale_ex(np.array([1.0, 2.0]), 0, 3, mid=1.0)   # mid - x_min == 0


def test_rescale_nan_makes_both_extremes_nan(oracle):
    """np.min and np.max of a float array return its first NaN under numba
    (numba 0.54.1 numba/np/arraymath.py:516-531, 599-614), so one NaN of
    either sign makes every value that a form rescales NaN, values outside a
    one-sided mode's half are copied, and power_rescale's bounds start at
    NaN and stay there (util.py:125-126, 191-195)."""
    for nan in (np.nan, -np.nan):
        x = np.array([1.0, nan, -2.0, 3.0])
        assert np.isnan(oracle.rescale(x, 0, 1)).all()
        assert np.isnan(oracle.rescale_ex(x, -1.0, 1.0)).all()
        assert np.isnan(oracle.rescale_ex(x, -1.0, 1.0, mid=0.0)).all()
        lo = oracle.rescale_ex(x, -1.0, 1.0, mid=0.0, mode="lower")
        assert lo[0] == 1.0 and np.isnan(lo[1]) and np.isnan(lo[2]) and lo[3] == 3.0
        up = oracle.rescale_ex(x, -1.0, 1.0, mid=0.0, mode="upper")
        assert np.isnan(up[0]) and np.isnan(up[1]) and up[2] == -2.0 and np.isnan(up[3])
        m = np.array([True, False, True, False])
        pr = oracle.power_rescale(x, m, 1, 2.0)
        assert np.isnan(pr[0]) and np.isnan(pr[2]) and pr[3] == 3.0
    assert np.signbit(np.array([-np.nan]))[0]   # the second case is the negative NaN


def test_mask_is_inclusive(oracle):
    m = oracle.make_bool_elevation_mask(np.array([-1.0, 0.0, 0.5, 1.0]), 0.5)
    assert m.tolist() == [True, True, True, False]


def test_find_percent_val(oracle):
    assert oracle.find_percent_val(-4000, 8850, 55.0) == -4000 + 12850 * 55.0 / 100.0
    assert oracle.find_percent_val(0, 10, 0.0) == 5.0       # outside (0, 100): 50%
    assert oracle.find_percent_val(0, 10, 100.0) == 5.0


def test_power_rescale_hand_worked(oracle):
    x = np.array([1.0, 9.0, 4.0, 20.0, 0.0])
    mask = np.array([1, 1, 1, 0, 0], np.uint8)
    # mode 1, masked values 1, 9, 4: range [1, 9]; t = (x-1)/8, t**0.5, back
    out = oracle.power_rescale(x, mask, 1, 0.5)
    t = np.sqrt(np.array([0.0, 1.0, 3.0 / 8.0]))
    assert np.allclose(out[:3], t * 8 + 1, rtol=0, atol=1e-12)
    assert (out[3:] == x[3:]).all()
    # mode 0 rescales the unmasked values 20, 0: range [0, 20], power 2
    out0 = oracle.power_rescale(x, mask, 0, 2.0)
    assert (out0[:3] == x[:3]).all() and out0[3] == 20.0 and out0[4] == 0.0


def test_power_rescale_sequential_quirk(oracle):
    """Masked values 7 then 5: both are new running minima (from x_max), so
    the mask maximum keeps its start value x_min = 0 and the range is -5."""
    x = np.array([7.0, 5.0, 10.0, 0.0])
    mask = np.array([1, 1, 0, 0], np.uint8)
    out = oracle.power_rescale(x, mask, 1, 2.0)
    # t = (x - 5) / -5 -> -0.4, 0; t**2 -> 0.16, 0; back: t * -5 + 5
    assert np.allclose(out, [0.16 * -5 + 5, 5.0, 10.0, 0.0], rtol=0, atol=1e-12)


def test_height_chain_bounds_and_ocean(oracle):
    rng = np.random.default_rng(0)
    h = rng.uniform(-1, 1, 5000)
    out, mask, level = oracle.height_chain(h, -4000, 8850, 55.0)
    assert out.min() == -4000 and out.max() == 8850      # split rescale endpoints
    assert level == -4000 + 12850 * 55.0 / 100.0
    # ocean vertices end below (or at) 0, land above, after the shift and mid=0 rescale
    assert (out[mask] <= 0).all() and (out[~mask] > 0).all()
