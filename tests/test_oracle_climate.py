"""Pin the oracle's insolation (climate.py:193-201, 416-590) on the CPU.

* The equator row of a day at zero tilt is sum_k max(cos(k deg), 0) over
  k = -180..179, which is the Dirichlet sum sin(89.5 deg) / sin(0.5 deg);
  every latitude gets that times cos(lat) (the max(cos|lat|, 0) factor),
  north and south alike.
* At 90 degrees of tilt the north pole faces the sun all day: exactly 360.
* One step at rotation -90 turns (0, 1, 0) to the sub-solar point: 1; a sign
  slip in the rotation turns it away (0).
* Interpolation through a table that is linear in latitude returns the
  latitude itself, which pins the negative-latitude layout (181 + lat).
* A year at zero tilt is 360 identical days.
All with float32 accumulation, so sums of 360 terms are held to 2e-3.
"""
import math

import numpy as np


def _ll(lat, lon=0.0, r=1.0):
    d = math.pi / 180
    return np.array([r * math.cos(lat * d) * math.cos(lon * d),
                     r * math.cos(lat * d) * math.sin(lon * d), r * math.sin(lat * d)])


def test_day_at_zero_tilt_closed_form(oracle):
    t = oracle.insolation_slice(1.0, 0.0)
    assert t.dtype == np.float32 and t.shape == (181,)
    eq = math.sin(math.radians(89.5)) / math.sin(math.radians(0.5))
    assert abs(t[0] - eq) < 2e-3
    for lat in (10, 33, 60, 89):
        assert abs(t[lat] - eq * math.cos(math.radians(lat))) < 2e-3
        assert t[lat] == t[181 - lat]                      # latitude -lat at 181 - lat
    assert t[90] < 1e-10 and t[91] < 1e-10


def test_pole_faces_the_sun_at_90_degrees_tilt(oracle):
    t = oracle.insolation_slice(1.0, 90.0)
    assert t[90] == 360.0 and t[91 + 89] < 1e-10 and t[-90 + 181] == 0.0


def test_rotation_signs(oracle):
    arr = np.zeros(2, np.float32)
    v = np.array([[0.0, 1.0, 0.0], [0.0, -1.0, 0.0]])
    oracle.sample_insolation(arr, v, 1.0, -90.0, 0.0)
    assert arr[0] == 1.0 and arr[1] == 0.0
    arr[:] = 0
    oracle.sample_insolation(arr, v, 1.0, 90.0, 0.0)
    assert arr[0] == 0.0 and arr[1] == 1.0


def test_interpolation_through_a_linear_table(oracle):
    table = np.zeros(181, np.float32)
    table[:91] = np.arange(91)
    table[91:] = np.arange(-90, 0)
    lats = [-89.5, -45.25, -0.75, 0.0, 12.5, 60.125, 89.0]
    verts = np.array([_ll(a, 17.0 * i) for i, a in enumerate(lats)])
    got = oracle.interpolate_insolation(verts, table, 1.0)
    assert np.allclose(got, lats, atol=1e-5, rtol=0)


def test_brute_day_equals_the_slice_on_the_equator(oracle):
    t = oracle.insolation_slice(1.0, 0.0)
    d = oracle.brute_daily_insolation(np.array([_ll(0.0), _ll(30.0)]), 1.0, 0.0)
    assert d[0] == t[0] and abs(d[1] - t[30]) < 2e-3


def test_year_at_zero_tilt_is_360_days(oracle):
    verts = np.array([_ll(a, 33.0) for a in (-70.0, -12.5, 0.0, 41.3)])
    y = oracle.yearly_insolation(verts, 1.0, 0.0)
    day = oracle.interpolate_insolation(verts, oracle.insolation_slice(1.0, 0.0), 1.0)
    assert np.allclose(y, 360.0 * day, rtol=1e-5, atol=0)


def test_surface_temp_closed_forms(oracle):
    """climate.py:351-412: max(cos(|lat - tilt| deg), 0) whatever the
    altitudes (their weight is 0): 1 where the latitude equals the tilt,
    cos(lat - tilt) elsewhere, 0 beyond 90 degrees from it, north and south
    mirrored about the tilt."""
    lats = [-80.0, -30.0, 0.0, 23.44, 45.0, 89.0]
    verts = np.array([_ll(a, 11.0 * i) for i, a in enumerate(lats)])
    for alts in (np.linspace(-5, 5, len(lats)), np.array([1e4, -3, 0, 7, 2, 1e-3])):
        t = oracle.assign_surface_temp(verts, alts, 1.0, 23.44)
        assert t.dtype == np.float32
        for a, v in zip(lats, t):
            assert abs(v - max(math.cos(math.radians(abs(a - 23.44))), 0.0)) < 1e-6
    # exactly at the tilt and exactly 90 degrees away
    t = oracle.assign_surface_temp(np.array([_ll(30.0), _ll(-60.0), _ll(-70.0)]),
                                   np.array([0.0, 1.0, 2.0]), 1.0, 30.0)
    assert abs(t[0] - 1.0) < 1e-6 and 0 <= t[1] < 1e-6 and t[2] == 0.0


def test_surface_temp_constant_altitudes_raise(oracle):
    import pytest
    with pytest.raises(ZeroDivisionError):
        oracle.assign_surface_temp(np.array([_ll(0.0), _ll(10.0)]), np.zeros(2), 1.0, 0.0)
