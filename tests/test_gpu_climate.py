"""GPU parity of the insolation (climate.py:193-201, 416-590) against the
oracle.

The per-angle sines and cosines come from the same C library as the
oracle's; the per-vertex asin, atan2 and cos are the device's, which can
differ from glibc's in the last bit.  Each float32 accumulation step then
rounds to the same float nearly always and otherwise to a neighbour, so the
bound is one float32 ulp of the result per summed step:
|gpu - oracle| <= n_steps * ulp32(max |oracle|).  Measured agreement is
printed.
"""
import numpy as np
import pytest
import torch

from nixis_amd import climate, util

pytestmark = pytest.mark.gpu


def _bound(ref, nsteps):
    return nsteps * float(np.spacing(np.float32(np.abs(ref).max())))


def _report(name, got, ref, nsteps):
    err = float(np.abs(got.astype(np.float64) - ref.astype(np.float64)).max())
    tol = _bound(ref, nsteps)
    print(f"{name}: max |gpu - oracle| {err:.3e} (bound {tol:.3e}), exact {np.mean(got == ref):.4f}",
          flush=True)
    assert got.dtype == np.float32 and err <= tol


@pytest.mark.parametrize("tilt", [0.0, 23.44, -61.0])
def test_sample_insolation_matches_oracle(oracle, tilt):
    pts, _ = util.create_mesh(64)
    for rot, n in ((0.0, 1), (-180.0, 360)):
        got = np.zeros(pts.shape[0], np.float32)
        if n == 1:
            climate.sample_insolation(got, pts, 1.0, rot, tilt)
        else:
            got = climate.brute_daily_insolation(pts, None, 1.0, tilt)
        ref = oracle.sample_insolation(np.zeros(pts.shape[0], np.float32), pts, 1.0, rot, tilt,
                                       n, 1.0)
        _report(f"tilt {tilt} steps {n}", got, ref, n)


def test_slice_and_daily_match_oracle(oracle):
    pts, _ = util.create_mesh(48)
    for tilt in (0.0, 12.5, 90.0):
        t = climate.calc_insolation_slice(1.0, tilt)
        rt = oracle.insolation_slice(1.0, tilt)
        _report(f"slice tilt {tilt}", t, rt, 360)
        d = climate.calc_daily_insolation(pts, None, 1.0, tilt)
        rd = oracle.interpolate_insolation(pts, rt, 1.0)
        _report(f"daily tilt {tilt}", d, rd, 362)
    assert climate.calc_insolation_slice(1.0, 90.0)[90] == 360.0   # the pole faces the sun


def test_interpolation_through_a_linear_table():
    table = np.zeros(181, np.float32)
    table[:91] = np.arange(91)
    table[91:] = np.arange(-90, 0)
    lats = np.array([-89.5, -45.25, -0.75, 0.0, 12.5, 60.125, 89.0])
    d = np.pi / 180
    verts = np.stack([np.cos(lats * d), np.zeros_like(lats), np.sin(lats * d)], axis=1)
    out = np.empty(lats.size, np.float32)
    climate.interpolate_insolation(verts, table, out, 1.0)
    assert np.allclose(out, lats, atol=1e-5, rtol=0)


def test_yearly_matches_oracle(oracle):
    pts, _ = util.create_mesh(24)
    for tilt in (0.0, 23.44):
        got = climate.calc_yearly_insolation(pts, None, 1.0, tilt)
        ref = oracle.yearly_insolation(pts, 1.0, tilt)
        # 360 days, each a table entry within 360 ulp of the daily magnitude
        # plus one rounding of the interpolation, then 360 float32 additions
        day_ulp = float(np.spacing(np.float32(ref.max() / 360)))
        tol = 360 * 361 * day_ulp + 360 * float(np.spacing(np.float32(ref.max())))
        err = float(np.abs(got.astype(np.float64) - ref).max())
        print(f"yearly tilt {tilt}: max |gpu - oracle| {err:.3e} (bound {tol:.3e}), "
              f"exact {np.mean(got == ref):.4f}", flush=True)
        assert got.dtype == np.float32 and err <= tol


def test_constant_altitudes_raise_as_the_reference():
    pts, _ = util.create_mesh(4)
    with pytest.raises(ZeroDivisionError):
        climate.calc_daily_insolation(pts, np.zeros(pts.shape[0]), 1.0, 0.0)


def test_yearly_full_size_properties():
    """d=500 (2,500,002 vertices): a year at 23.44 degrees of tilt is
    symmetric north/south (the seasons mirror), bounded by 360 days of the
    equator's daily maximum, and largest at the equator."""
    pts, _ = util.create_mesh_device(500, cells=False)
    y = climate.yearly_insolation_device(pts, 1.0, 23.44)
    mirror = pts.clone()
    mirror[:, 2] = -mirror[:, 2]
    ym = climate.yearly_insolation_device(mirror, 1.0, 23.44)
    torch.cuda.synchronize()
    eq_day = climate.calc_insolation_slice(1.0, 0.0)[0]
    y, ym = y.cpu().numpy(), ym.cpu().numpy()
    assert np.abs(y - ym).max() <= 1e-3 * y.max()
    assert (y >= 0).all() and y.max() <= 360 * eq_day * (1 + 1e-5)
    z = pts[:, 2].cpu().numpy()
    assert y[np.abs(z) < 0.01].mean() > y[np.abs(z) > 0.99].mean()


@pytest.mark.parametrize("tilt", [0.0, 23.44, -61.0])
def test_surface_temp_matches_oracle(oracle, tilt):
    """climate.py:351-412: the device's asin and cos against glibc's, so one
    float32 ulp of the result; values in [0, 1]."""
    pts, _ = util.create_mesh(64)
    alts = np.linspace(-9000, 9000, pts.shape[0])
    got = climate.assign_surface_temp(pts, alts, 1.0, tilt)
    ref = oracle.assign_surface_temp(pts, alts, 1.0, tilt)
    _report(f"surface temp tilt {tilt}", got, ref, 1)
    assert got.min() >= 0 and got.max() <= 1
    with pytest.raises(ZeroDivisionError):
        climate.assign_surface_temp(pts, np.full(pts.shape[0], 3.0), 1.0, tilt)
    dev = climate.assign_surface_temp_device(torch.from_numpy(pts).cuda(), 1.0, tilt)
    assert np.array_equal(dev.cpu().numpy(), got)
