"""Pin the FP64 oracle's init / noise3d / sample_octaves / rescale (CPU only).

The reference ships no tests or golden vectors (its numba path cannot run
here), so the oracle is pinned by: a second, structurally different
derivation of noise3d (tests/noise_offsets.py) on random and crafted points,
the hash lattice's period, value bounds, table invariants of init, and the
closed forms of the octave weighting and rescale.
"""
import numpy as np
import pytest

import noise_offsets as NOFF

SEEDS = [0, 1, 7, 99, 1234, -1, 2**31 - 1, 2**31, 2**63 - 1, -(2**63)]


@pytest.mark.parametrize("seed", SEEDS)
def test_init_tables_structure(oracle, seed):
    perm, pgi = oracle.init(seed)
    assert sorted(perm.tolist()) == list(range(256))       # a permutation
    assert (pgi == 3 * (perm % 24)).all()                   # opensimplex.py:104
    p2, g2 = oracle.init(seed)
    assert (p2 == perm).all() and (g2 == pgi).all()         # deterministic


def test_init_lcg_reading(oracle):
    """The int32 truncation of over() (opensimplex.py:110-112): the state after
    each step fits int32, and distinct seeds that agree mod 2^32 after the
    first int64 step give identical tables."""
    C1, C2 = 6364136223846793005, 1442695040888963407

    def step(s):
        t = (s * C1 + C2) & 0xFFFFFFFF          # low 32 bits of the int64 result
        return t - (1 << 32) if t >= (1 << 31) else t

    # A seed and seed + 2^32 have the same low 32 bits of s*C1+C2 -> same tables.
    a = oracle.init(12345)
    b = oracle.init(12345 + (1 << 32))
    assert (a[0] == b[0]).all()
    # Python big-int version of the LCG and shuffle (host-language check of
    # the C restatement's integer semantics, incl. the floored modulo).
    for seed in (0, 5, -77, 2**40 + 3):
        s = step(step(step(seed)))
        source = list(range(256))
        perm = [0] * 256
        for i in range(255, -1, -1):
            s = step(s)
            r = (s + 31) % (i + 1)
            perm[i] = source[r]
            source[r] = source[i]
        assert perm == oracle.init(seed)[0].tolist()


def _points(rng, n, scale):
    return rng.uniform(-scale, scale, size=(n, 3))


def _crafted():
    """Lattice faces, region boundaries, score ties and large coordinates."""
    pts = []
    # points whose skewed coordinates are integers: x = xs - so with
    # so = (xs+ys+zs)*(-1/6) / (1 + 3*(-1/6)) -> x = xs + (xs+ys+zs)/3
    for xs in (-3, 0, 1, 7):
        for ys in (-2, 0, 5):
            for zs in (0, 1, -4):
                t = (xs + ys + zs) / 3.0
                pts.append((xs + t, ys + t, zs + t))
    # in_sum exactly 1 and 2 inside the cell at the origin: skewed (a,b,c)
    for a, b, c in [(0.5, 0.25, 0.25), (0.75, 0.75, 0.5), (1 / 3, 1 / 3, 1 / 3),
                    (0.5, 0.5, 0.0), (0.25, 0.25, 0.25), (0.6, 0.6, 0.6),
                    (0.9, 0.9, 0.2), (0.1, 0.45, 0.45), (0.5, 0.5, 0.5)]:
        t = (a + b + c) / 3.0
        pts.append((a + t, b + t, c + t))
    # large coordinates (high octaves: frequency up to 2^16)
    for s in (2.0**10, 2.0**14, 2.0**16):
        pts += [(s * 0.3, -s * 0.7, s * 0.11), (s, s, s), (-s, 0.5, s / 3)]
    return np.array(pts, dtype=np.float64)


@pytest.mark.parametrize("seed,scale,n", [(1234, 4.0, 3000), (7, 100.0, 3000),
                                          (99, 3e4, 2000), (0, 1.0, 2000)])
def test_noise3d_matches_offset_derivation(oracle, seed, scale, n):
    perm, pgi = oracle.init(seed)
    rng = np.random.default_rng(seed + 17)
    pts = np.vstack([_points(rng, n, scale), _crafted()])
    for p in pts:
        v, sel = oracle.noise3d(*p, perm, pgi, with_sel=True)
        w, sel2 = NOFF.noise3d(*p, perm, pgi, with_sel=True)
        assert (sel.astype(np.int64) == sel2).all(), (p, sel, sel2)
        assert abs(v - w) <= 1e-13, (p, v, w)


def test_noise3d_regions_all_exercised(oracle):
    """The random sample covers the three regions and both extra-vertex
    cases of each (so the derivation check above is not vacuous)."""
    perm, pgi = oracle.init(3)
    rng = np.random.default_rng(5)
    codes = set()
    for p in _points(rng, 4000, 10.0):
        _, sel = oracle.noise3d(*p, perm, pgi, with_sel=True)
        codes.add(int(sel[3]))
    regions = {c & 3 for c in codes}
    assert regions == {0, 1, 2}
    # extra-vertex layouts per region: 3 + 3 in each tetrahedron, up to 15
    # in the octahedron
    for r in (0, 1, 2):
        assert len({c for c in codes if c & 3 == r}) >= 6


def test_noise3d_period_512(oracle):
    """Skewed lattice shifts by 256 (hash period) for a (512,512,512) offset."""
    perm, pgi = oracle.init(42)
    rng = np.random.default_rng(1)
    for p in _points(rng, 500, 50.0):
        a = oracle.noise3d(*p, perm, pgi)
        b = oracle.noise3d(*(p + 512.0), perm, pgi)
        assert abs(a - b) < 1e-9


def test_noise3d_bounded_and_nonconstant(oracle):
    perm, pgi = oracle.init(11)
    rng = np.random.default_rng(2)
    pts = _points(rng, 20000, 20.0)
    v = oracle.noisearr3d(pts[:, 0], pts[:, 1], pts[:, 2], perm, pgi)
    assert np.abs(v).max() <= 1.0
    assert v.std() > 0.1 and abs(v.mean()) < 0.05


def test_sample_octaves_is_sum_of_sample_noise(oracle):
    """terrain.py:40-45: octaves = running sum of sample_noise at f*=lac, a*=pers."""
    perm, pgi = oracle.init(1234)
    rng = np.random.default_rng(3)
    v = rng.normal(size=(500, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    R = 2.5
    verts = v * R
    acc = np.zeros(500)
    f, a = 1.5, 0.4
    for _ in range(5):
        acc += oracle.sample_noise(verts, perm, pgi, f / R, a / R, R)
        f *= 2.0
        a *= 0.5
    e = oracle.sample_octaves(verts, None, perm, pgi, 5, 1.5, 0.4, 2.0, 0.5, R)
    assert (e == acc).all()
    # and each sample_noise is ((noise + 1) * 0.5 * s) * r of noise at v * f
    out, raw = oracle.sample_noise(verts, perm, pgi, 1.5, 0.7, 3.0, raw=True)
    for k in range(0, 500, 50):
        n = oracle.noise3d(*(verts[k] * 1.5), perm, pgi)
        assert raw[k] == n
        assert out[k] == (n + 1) * 0.5 * 0.7 * 3.0


def test_rescale_closed_form(oracle):
    x = np.array([3.0, -1.0, 7.0, 2.5])
    y = oracle.rescale(x, -9000, 9000)
    assert y.min() == -9000 and y.max() == 9000
    np.testing.assert_allclose(y, (x + 1.0) / 8.0 * 18000 - 9000, rtol=0, atol=1e-9)
    with pytest.raises(ZeroDivisionError):
        oracle.rescale(np.ones(4), 0, 1)
