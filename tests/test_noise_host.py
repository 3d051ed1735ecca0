"""The product's noise evaluation (nixis_amd/csrc/nx_noise_dev.h) compiled for
the host and checked against the FP64 oracle on the CPU.

The header is written to compile for the host as well as for gfx950 (its
region logic works on lane masks, which reduce to one lane here), so the
same source that the kernels run can be pinned without a GPU: the lattice
cell, region, extra vertices and gradient hashes of every point equal the
oracle's selection record, and the FP32 value is within 1e-5 of noise3d.
The points cover random coordinates at several scales, lattice faces, the
in_sum = 1 and in_sum = 2 boundaries and exact score ties (skewed in-cell
coordinates on a 1/8 grid), which is where a `<` written for a `<=` in the
region logic changes the selection.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from test_oracle_noise import _crafted

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NOISE_TOL = 1e-5


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("noise_host") / "libnoisehost.so")
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "nixis_amd", "csrc"),
           os.path.join(ROOT, "tests", "noise_host.cpp"), "-o", so]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lib = ctypes.CDLL(so)
    lib.host_noise3.restype = ctypes.c_int64
    lib.host_noise3.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_void_p]

    def run(pts, perm, pgi):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        perm = np.ascontiguousarray(perm, dtype=np.int32)
        pgi = np.ascontiguousarray(pgi, dtype=np.int32)
        val = np.empty(pts.shape[0], np.float32)
        sel = np.empty((pts.shape[0], 5), np.int32)
        bad = lib.host_noise3(pts.ctypes.data, pts.shape[0], perm.ctypes.data, pgi.ctypes.data,
                              val.ctypes.data, sel.ctypes.data)
        assert bad == 0
        return val, sel
    return run


def _oracle_sel(oracle, pts, perm, pgi):
    """noise3d and its selection record at each point: one octave of
    frequency 1 and radius 1 is noise3d at the point itself."""
    _, raw, sel = oracle.sample_octaves(pts, None, perm, pgi, 1, 1.0, 0.4, 2.0, 0.5, 1.0,
                                        raw=True, sel=True)
    return raw[0], sel[0]


def _tie_points(rng, n):
    """Points whose skewed coordinates have in-cell parts on a 1/8 grid:
    exact ties between scores, in_sum exactly 1 or 2, p = 1 sums."""
    cell = rng.integers(-40, 40, size=(n, 3)).astype(np.float64)
    frac = rng.integers(0, 8, size=(n, 3)) / 8.0
    s = cell + frac
    t = s.sum(axis=1, keepdims=True) / 3.0      # unskew: x = xs + (xs+ys+zs)/3
    return s + t


@pytest.mark.parametrize("seed,scale", [(1234, 4.0), (7, 300.0), (99, 5e4), (0, 1.0)])
def test_host_build_matches_oracle(oracle, host, seed, scale):
    rng = np.random.default_rng(seed)
    pts = np.concatenate([rng.uniform(-scale, scale, size=(60_000, 3)), _crafted(),
                          _tie_points(rng, 60_000)])
    perm, pgi = oracle.init(seed)
    ref, rsel = _oracle_sel(oracle, pts, perm, pgi)
    val, sel = host(pts, perm, pgi)
    mism = int((sel != rsel).any(axis=1).sum())
    assert mism == 0, f"{mism} of {len(pts)} points select differently"
    err = float(np.abs(val.astype(np.float64) - ref).max())
    print(f"\nseed {seed} scale {scale}: {len(pts)} points, max |host - fp64| {err:.2e}")
    assert err <= NOISE_TOL
    # every region, and 24 distinct region / extra-vertex layouts
    regions = rsel[:, 3] & 3
    assert set(np.unique(regions)) == {0, 1, 2}
    assert len(np.unique(rsel[:, 3])) >= 24


@pytest.mark.parametrize("old,new", [
    ("zx_lt = lanes(zins < xins)", "zx_lt = lanes(zins <= xins)"),
    ("ab_le = lanes(asC <= bsC)", "ab_le = lanes(asC < bsC)"),
    ("inB = lanes(in_sum >= 2)", "inB = lanes(in_sum > 2)"),
    ("vz = lanes(wB < zins)", "vz = lanes(wB <= zins)"),
])
def test_points_catch_a_slip_in_the_region_logic(oracle, tmp_path, old, new):
    """The same points run through a copy of the header with one comparison
    changed at its tie (built in a temporary directory) select differently
    somewhere: the tie points reach every boundary the test relies on."""
    src = open(os.path.join(ROOT, "nixis_amd", "csrc", "nx_noise_dev.h")).read()
    assert src.count(old) == 1
    hdr = tmp_path / "nx_noise_dev.h"
    hdr.write_text(src.replace(old, new))
    so = str(tmp_path / "libmut.so")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                        "-I", str(tmp_path), "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "noise_host.cpp"), "-o", so],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lib = ctypes.CDLL(so)
    lib.host_noise3.restype = ctypes.c_int64
    rng = np.random.default_rng(11)
    pts = np.ascontiguousarray(np.concatenate([_crafted(), _tie_points(rng, 60_000)]))
    perm, pgi = oracle.init(11)
    _, rsel = _oracle_sel(oracle, pts, perm, pgi)
    val = np.empty(pts.shape[0], np.float32)
    sel = np.empty((pts.shape[0], 5), np.int32)
    lib.host_noise3(ctypes.c_void_p(pts.ctypes.data), ctypes.c_int64(pts.shape[0]),
                    ctypes.c_void_p(perm.ctypes.data), ctypes.c_void_p(pgi.ctypes.data),
                    ctypes.c_void_p(val.ctypes.data), ctypes.c_void_p(sel.ctypes.data))
    assert (sel != rsel).any(axis=1).sum() > 0
