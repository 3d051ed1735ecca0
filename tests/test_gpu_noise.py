"""GPU parity of the fused noise kernel against the FP64 oracle.

Bar: lattice cell, region, extra vertices and gradient hashes bit-exact
(dbg_sel == oracle selection); per-octave noise within 1e-5 absolute of the
FP64 value; elevations within sum_o(0.5 * strength_o * radius) * 1e-5.
"""
import numpy as np
import pytest
import torch

from nixis_amd import _lib, opensimplex as osx, terrain, util

NOISE_TOL = 1e-5
pytestmark = pytest.mark.gpu
DEV = "cuda"


def _run(pts, seed, n_oct, f0, a0, lac, pers, R=1.0, dbg=True):
    perm, pgi = osx.init(seed)
    rough, strength = terrain.octave_scalars(n_oct, f0, a0, lac, pers, R)
    V = pts.shape[0]
    dv = torch.from_numpy(np.ascontiguousarray(pts)).to(DEV)
    dn = torch.empty((n_oct, V), dtype=torch.float32, device=DEV) if dbg else None
    ds = torch.empty((n_oct, V, 5), dtype=torch.int32, device=DEV) if dbg else None
    e = terrain.sample_octaves_device(dv, None, perm, pgi, rough, strength, R, dn, ds)
    _lib.Context.get(dv.device).sync_check()
    out = e.cpu().numpy()
    if dbg:
        return out, dn.cpu().numpy(), ds.cpu().numpy(), strength
    return out, None, None, strength


def _check(oracle, pts, seed, n_oct, f0, a0, lac, pers, R=1.0, label=""):
    e, dn, ds, strength = _run(pts, seed, n_oct, f0, a0, lac, pers, R)
    perm, pgi = oracle.init(seed)
    e_ref, raw, sel = oracle.sample_octaves(pts, None, perm, pgi, n_oct, f0, a0, lac, pers, R,
                                            raw=True, sel=True)
    mism = int((ds != sel).any(axis=2).sum())
    assert mism == 0, f"{label}: {mism} vertex-octaves with a different lattice selection"
    per_oct = np.abs(dn.astype(np.float64) - raw).max(axis=1)
    print(f"\n{label} per-octave max |noise - fp64|: " +
          " ".join(f"{x:.2e}" for x in per_oct))
    assert per_oct.max() <= NOISE_TOL
    bound = float(np.sum(0.5 * strength * R)) * NOISE_TOL
    assert np.abs(e - e_ref).max() <= bound
    return per_oct


def test_noise_small_planet(oracle):
    pts = util.create_points(32)
    _check(oracle, pts, 1234, 8, 1.5, 0.4, 2.0, 0.5, label="d=32 8oct")


@pytest.mark.parametrize("seed", [0, -1, 99, 2**31 + 5])
def test_noise_seeds_and_world_radius(oracle, seed):
    pts = util.create_points(20) * 6378100.0
    _check(oracle, pts, seed, 7, 1.5, 0.4, 2.5, 0.5, R=6378100.0, label=f"seed {seed} R=earth")


def test_noise_config_plain_small(oracle):
    """d=200, 8 octaves, f0 1.5, lacunarity 2, persistence 0.5, seed 1234."""
    pts = util.create_points(200)
    _check(oracle, pts, 1234, 8, 1.5, 0.4, 2.0, 0.5, label="c1 d=200")


def test_noise_at_scale_subsample(oracle):
    """d=2000 / 16 octaves / f0 2.0 / persistence 0.55 / seed 7: the full
    planet runs without debug output; a strided subsample of its vertices is
    checked in full against the oracle (bit-exact selection, per-octave
    error), and the full run must equal the debug run on that subsample."""
    pts = util.create_points(2000)
    sub = pts[::97].copy()
    per_oct = _check(oracle, sub, 7, 16, 2.0, 0.4, 2.0, 0.55, label="c2 d=2000 subsample")
    assert per_oct.shape == (16,)
    full, _, _, _ = _run(pts, 7, 16, 2.0, 0.4, 2.0, 0.55, dbg=False)
    part, _, _, _ = _run(sub, 7, 16, 2.0, 0.4, 2.0, 0.55, dbg=False)
    assert (full[::97] == part).all()
    assert np.isfinite(full).all()


def test_noise_crafted_points(oracle):
    """Lattice faces, region boundaries and frequencies up to 2^16."""
    import noise_offsets  # noqa: F401  (same crafted set as the oracle tests)
    from test_oracle_noise import _crafted
    pts = _crafted()
    perm, pgi = osx.init(1234)
    ref = oracle.noisearr3d(pts[:, 0], pts[:, 1], pts[:, 2], perm, pgi)
    got = osx.noisearr3d(pts[:, 0], pts[:, 1], pts[:, 2], perm, pgi)
    assert np.abs(got - ref).max() <= NOISE_TOL
    assert abs(osx.noise3d(0.3, -0.2, 0.9, perm, pgi) -
               oracle.noise3d(0.3, -0.2, 0.9, perm, pgi)) <= NOISE_TOL


def test_noise_selection_on_ties_and_crafted_points(oracle):
    """Points whose in-cell coordinates tie (a 1/8 grid: equal scores,
    in_sum exactly 1 or 2, pair sums exactly 1) and the crafted lattice-face
    and region-boundary points: the device build selects the oracle's
    lattice points octave by octave, on the doubling path (lacunarity 2,
    frequencies 1, 2, 4, 8) and on the direct one (lacunarity 3); scaling by
    2 or 3 keeps the in-cell parts on the 1/8 grid, so every octave has ties."""
    from test_noise_host import _tie_points
    from test_oracle_noise import _crafted
    rng = np.random.default_rng(21)
    pts = np.ascontiguousarray(np.concatenate([_crafted(), _tie_points(rng, 20_000)]))
    _check(oracle, pts, 1234, 4, 1.0, 0.4, 2.0, 0.5, label="ties, doubling")
    _check(oracle, pts, 99, 3, 1.0, 0.4, 3.0, 0.5, label="ties, lacunarity 3")


def test_noise_accumulates_into_elevations(oracle):
    pts = util.create_points(16)
    perm, pgi = osx.init(5)
    base = np.linspace(-1, 1, pts.shape[0])
    e = terrain.sample_octaves(pts, base.copy(), perm, pgi, 4)
    ref = oracle.sample_octaves(pts, base.copy(), perm, pgi, 4)
    assert np.abs(e - ref).max() <= 4 * 0.5 * 0.4 * NOISE_TOL
    s = terrain.sample_noise(pts, perm, pgi, 2.0, 0.3, 1.0)
    s_ref = oracle.sample_noise(pts, perm, pgi, 2.0, 0.3, 1.0)
    assert np.abs(s - s_ref).max() <= 0.5 * 0.3 * NOISE_TOL


def test_noise_out_of_domain_raises():
    pts = np.array([[1.0, 0.0, 0.0], [2.0**40, 0.0, 0.0]])
    perm, pgi = osx.init(0)
    with pytest.raises(_lib.NixisRangeError):
        terrain.sample_octaves(pts, None, perm, pgi, 1)
    # the sticky error is cleared: a valid call afterwards succeeds
    terrain.sample_octaves(pts[:1], None, perm, pgi, 1)


def test_doubling_shortcut_is_bit_exact():
    """With frequencies that double (lacunarity 2) a wave doubles the
    previous octave's skewed coordinates, unless one of its vertices has a
    first-octave product in the subnormal range.  Interleaving such a vertex
    with every real one forces the direct path in every wave; selection,
    per-octave noise and elevations must be the same bit for bit."""
    pts = util.create_points(24)
    mixed = np.empty((2 * pts.shape[0], 3))
    mixed[0::2] = pts
    mixed[1::2] = [1e-310, 0.3, -0.2]
    a = _run(pts, 3, 12, 1.5, 0.4, 2.0, 0.5)
    b = _run(mixed, 3, 12, 1.5, 0.4, 2.0, 0.5)
    assert np.array_equal(a[0], b[0][0::2])
    assert np.array_equal(a[1], b[1][:, 0::2])
    assert np.array_equal(a[2], b[2][:, 0::2])


def test_domain_check_uses_the_highest_octave():
    """x = 2^16: octave 8 (frequency 1.5 * 2^7) stays inside |skewed| < 2^30,
    octave 16 (1.5 * 2^15) does not."""
    pts = np.array([[1.0, 0.0, 0.0], [2.0**16, 0.0, 0.0]])
    perm, pgi = osx.init(0)
    terrain.sample_octaves(pts, None, perm, pgi, 8)
    with pytest.raises(_lib.NixisRangeError):
        terrain.sample_octaves(pts, None, perm, pgi, 16)


def test_noise_empty_and_zero_octaves():
    perm, pgi = osx.init(0)
    assert terrain.sample_octaves(np.zeros((0, 3)), None, perm, pgi, 3).shape == (0,)
    e = terrain.sample_octaves(util.create_points(4), None, perm, pgi, 0)
    assert (e == 0).all()
