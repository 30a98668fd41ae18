"""CPU pins of the oracle's grids, coarse-grid spline, Planck function and
table readers.

Reference-held fixtures: code/grid.txt (im=190, L=2) and code/resfortm.txt
column 1 (im=64, L=1), copied under tests/golden/.  Where the reference's own
CPU sources build here (code/spline.cpp, code/read_tables.cpp, with the
driver oracle/ref_driver.cpp: `make -C oracle ref`), their outputs pin the
grid generator, interp3D and readT directly; those tests skip when the vectors
are absent.  Closed forms and scipy pin the rest.
"""
import math
import os

import numpy as np
import pytest

from helpers import GOLDEN, particle_table_dir, read_reference_tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_VECTORS = os.path.join(ROOT, "oracle", "_ref", "ref_vectors.txt")
SIGMA = 5.670373e-08


def _ref_vectors():
    if not os.path.exists(REF_VECTORS):
        pytest.skip("reference vectors not built (make -C oracle ref needs the reference tree)")
    out = {}
    with open(REF_VECTORS) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines):
        if not lines[i].strip():
            i += 1
            continue
        name, n = lines[i].split()
        n = int(n)
        out[name] = np.array([float(x) for x in lines[i + 1:i + 1 + n]])
        i += 1 + n
    return out


# ------------------------------------------------------------------ grids --
def test_ref_quad_grid_txt(oracle):
    g = np.loadtxt(os.path.join(GOLDEN, "grid.txt"))
    f, c = oracle.grid("ref_quad", 190, 2.0)
    idx = g[:, 0].astype(int)
    np.testing.assert_allclose(f[idx], g[:, 1], atol=6e-7)
    np.testing.assert_allclose(c[idx], g[:, 2], atol=6e-7)


def test_ref_quad_resfortm_centres(oracle):
    r = np.loadtxt(os.path.join(GOLDEN, "resfortm.txt"))
    f, c = oracle.grid("ref_quad", 64, 1.0)
    np.testing.assert_allclose(c[1:65], r[:, 0], rtol=0, atol=8e-8)   # printed to 8 digits


def test_grids_equal_reference_build(oracle):
    v = _ref_vectors()
    im0 = 64
    for lev in range(5):
        n = im0 >> lev
        f, c = oracle.grid("ref_quad", n, 1.0, n0=im0)
        np.testing.assert_array_equal(f, v["xu%d" % lev].astype(np.float32))
        np.testing.assert_array_equal(c, v["x%d" % lev].astype(np.float32))
        fu, cu = oracle.grid("uniform", n, 1.0)
        np.testing.assert_allclose(fu, v["yu%d" % lev], rtol=0, atol=2e-7)
        np.testing.assert_allclose(cu, v["y%d" % lev], rtol=0, atol=2e-7)
        np.testing.assert_allclose(fu, v["zu%d" % lev], rtol=0, atol=2e-7)


@pytest.mark.parametrize("n,beta", [(32, 1.5), (96, 1.5), (128, 2.0)])
def test_tanh_grid_closed_form(oracle, n, beta):
    L = 2.0
    f, c = oracle.grid("tanh", n, L, beta=beta)
    j = np.arange(n + 1)
    want = 0.5 * L * (1.0 + np.tanh(beta * (2.0 * j / n - 1.0)) / np.tanh(beta))
    np.testing.assert_allclose(f, want, atol=2e-7)
    assert f[0] == 0.0 and f[-1] == np.float32(L)
    np.testing.assert_allclose(f + f[::-1], L, atol=4e-7)          # symmetric about L/2
    assert np.all(np.diff(f) > 0)
    np.testing.assert_allclose(c[1:-1], 0.5 * (f[1:] + f[:-1]), rtol=1e-7)


# ----------------------------------------------------------------- spline --
def _natural_spline_3d(xf, yf, zf, T, xc, yc, zc):
    from scipy.interpolate import CubicSpline
    a = np.asarray(T, np.float64)                                   # [k, j, i]
    a = CubicSpline(xf, a, axis=2, bc_type="natural")(xc)
    a = CubicSpline(yf, a, axis=1, bc_type="natural")(yc)
    a = CubicSpline(zf, a, axis=0, bc_type="natural")(zc)
    return a


def test_interp3d_matches_scipy(oracle):
    rng = np.random.default_rng(7)
    nf, nc = (24, 20, 16), (12, 10, 8)
    fx, cx = oracle.grid("ref_quad", 24, 1.0)
    fy, cy = oracle.grid("tanh", 20, 2.0, beta=1.5)
    fz, cz = oracle.grid("uniform", 16, 1.0)
    gx, ccx = oracle.grid("ref_quad", 12, 1.0, n0=24)
    gy, ccy = fy[::2], None
    ccy = np.concatenate([[0.0], 0.5 * (gy[1:] + gy[:-1]), [0.0]]).astype(np.float32)
    gz, ccz = oracle.grid("uniform", 8, 1.0)
    T = (800 + 50 * rng.standard_normal((16, 20, 24))).astype(np.float32)
    got = oracle.interp3d(nf, cx, cy, cz, T.ravel(), nc, ccx, ccy, ccz).reshape(8, 10, 12)
    want = _natural_spline_3d(cx[1:-1], cy[1:-1], cz[1:-1], T, ccx[1:-1], ccy[1:-1], ccz[1:-1])
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=1e-4)


def test_interp3d_exact_for_linear_fields(oracle):
    nf, nc = (16, 12, 8), (8, 6, 4)
    fx, cx = oracle.grid("ref_quad", 16, 1.0)
    fy, cy = oracle.grid("tanh", 12, 2.0, beta=1.5)
    fz, cz = oracle.grid("uniform", 8, 3.0)
    gx, ccx = oracle.grid("ref_quad", 8, 1.0, n0=16)
    gy = fy[::2]
    ccy = np.concatenate([[0.0], 0.5 * (gy[1:] + gy[:-1]), [0.0]]).astype(np.float32)
    gz, ccz = oracle.grid("uniform", 4, 3.0)
    X, Y, Z = np.meshgrid(cx[1:-1], cy[1:-1], cz[1:-1], indexing="ij")
    T = (700 + 30 * X - 20 * Y + 5 * Z).transpose(2, 1, 0).astype(np.float32)
    got = oracle.interp3d(nf, cx, cy, cz, T.ravel(), nc, ccx, ccy, ccz).reshape(4, 6, 8)
    Xc, Yc, Zc = np.meshgrid(ccx[1:-1], ccy[1:-1], ccz[1:-1], indexing="ij")
    want = (700 + 30 * Xc - 20 * Yc + 5 * Zc).transpose(2, 1, 0)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-4)


def test_interp3d_matches_reference_build(oracle):
    """The reference's interp3D (FP32 spline/splint) against the oracle (FP64)."""
    v = _ref_vectors()
    n0 = 64
    grids = []
    for lev in range(5):
        n = n0 >> lev
        f, c = oracle.grid("ref_quad", n, 1.0, n0=n0)
        fu, cu = oracle.grid("uniform", n, 1.0)
        grids.append((n, c, cu))
    T0 = v["T0"].astype(np.float32)
    _, cx0, cu0 = grids[0]
    for lev in range(1, 5):
        n, cx, cu = grids[lev]
        got = oracle.interp3d((n0, n0, n0), cx0, cu0, cu0, T0, (n, n, n), cx, cu, cu)
        ref = v["T%d" % lev]
        # the reference interpolates in FP32: its own rounding is the tolerance
        np.testing.assert_allclose(got, ref, rtol=0, atol=3e-3)


# ----------------------------------------------------------------- Planck --
def test_planck_band_closed_form(oracle):
    C1, C2 = 3.741771790075259e-16, 0.014387741858429
    for T in (300.0, 573.0, 1000.0, 2000.0):
        for nu in (50.0, 500.0, 2500.0, 7000.0):
            want = C1 / math.pi * (100 * nu) ** 3 / math.expm1(C2 * 100 * nu / T)
            assert oracle.planck_band(T, nu) == pytest.approx(want, rel=2e-6)


def test_planck_band_integrates_to_sigma_t4(oracle):
    for T in (400.0, 1000.0, 1800.0):
        nu = np.arange(0.5, 40000.0, 1.0)
        s = sum(oracle.planck_band(T, float(x)) for x in nu[::10]) * 10.0 * 100.0
        assert s == pytest.approx(SIGMA * T ** 4 / math.pi, rel=2e-3)


def test_gauss_legendre_nodes():
    from dnsmc_amd import workloads
    x, w = workloads.gauss_legendre01(16)
    xr, wr = np.polynomial.legendre.leggauss(16)
    np.testing.assert_allclose(x, 0.5 * (xr + 1.0), atol=1e-12)
    np.testing.assert_allclose(w, 0.5 * wr, atol=1e-12)


# ----------------------------------------------------------------- tables --
def test_particle_tables_equal_reference_readT(oracle):
    v = _ref_vectors()
    t = read_reference_tables(particle_table_dir(), gas=False)
    np.testing.assert_array_equal(t["Tnb"], v["Tnb"].astype(np.float32))
    np.testing.assert_array_equal(t["kP"], v["kP"].astype(np.float32))
    np.testing.assert_array_equal(t["wvc"], v["wvc"].astype(np.float32))
    np.testing.assert_array_equal(t["kq"][:, 0, 0], v["kq"].astype(np.float32))
    np.testing.assert_array_equal(t["prob"].ravel(), v["prob"].astype(np.float32))


def test_kappamax_equals_reference_readT(oracle):
    """kappa at Tmax by linear interpolation of the Planck-mean table
    (code/read_tables.cpp:199-202), at the Tmax the driver used (750 K)."""
    v = _ref_vectors()
    t = read_reference_tables(particle_table_dir(), gas=False)
    ref = oracle.build_proposal(1, t, 750.0, 20000, np.ones(1))
    assert ref["kappamax"] == pytest.approx(float(v["kappamax"][0]), rel=1e-6)
