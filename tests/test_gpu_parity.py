"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.

Bars: bit-exact for integers (band/g draws, direction uniform bits, per-ray
step count, end type and visited-cell hash); FP32 per-ray values (D, G) within
2e-5 relative (only exp rounding differs: v_exp_f32 vs glibc expf); FP64 cell
tallies E, A within 1e-5 relative and Q = E - A within 1e-5 * E absolute.
"""
import math
import os

import numpy as np
import pytest

import dnsmc_amd
from dnsmc_amd import workloads
from helpers import legacy_conf_text, oracle_problem_dict, particle_table_dir, read_reference_tables, ref_subsequence

pytestmark = pytest.mark.gpu


def small_channel(n=(8, 12, 8), R=64, spectral=3, Tw=300.0, **kw):
    return workloads.channel_config(n, rays_per_cell=R, spectral=spectral, Tw=Tw, **kw)


def _decode(rays, nQ):
    band = (rays["w0"] >> 24) | (((rays["w1"] >> 24) & 3) << 8)
    g = rays["w1"] >> 26
    return band.astype(np.int64), g.astype(np.int64)


def _check_rays_and_tallies(oracle, S, cfg, gen, R):
    d = oracle_problem_dict(S, cfg)
    P = oracle.problem(d)
    nemit = S.nemit
    rays = S.rays(nemit * R)
    rec = S.records(nemit * R)
    assert len(rays) == nemit * R
    res = S.results(0)
    rel = []
    for cell in range(nemit):
        ref = gen(P, cell)
        got = rays[cell * R:(cell + 1) * R]
        np.testing.assert_array_equal(got["w0"], ref["w0"])
        np.testing.assert_array_equal(got["w1"], ref["w1"])
        np.testing.assert_allclose(got["w"], ref["w"], rtol=2e-7)
        np.testing.assert_allclose(got["winv"], ref["winv"], rtol=2e-7)
        i, j, k = cell % cfg.n[0] + 1, (cell // cfg.n[0]) % cfg.n[1] + 1, cell // (cfg.n[0] * cfg.n[1]) + 1
        band, g = _decode(ref, d["nQ"])
        for r in range(R):
            ut = (float(ref["w0"][r] & 0xFFFFFF) + d["u_off"]) * 2.0 ** -24
            up = (float(ref["w1"][r] & 0xFFFFFF) + d["u_off"]) * 2.0 ** -24
            o = oracle.trace_ray(P, i, j, k, ut, up, int(band[r]), int(g[r]))
            q = cell * R + r
            assert rec["steps"][q] == o.steps, (cell, r)
            assert rec["end"][q] == o.end_type, (cell, r)
            assert rec["hash"][q] == o.hash, (cell, r)
            assert abs(rec["D"][q] - o.D) <= 2e-5 * max(abs(o.D), 1e-30)
            assert abs(rec["G"][q] - o.G) <= 2e-5 * max(abs(o.G), 1e-30)
        E, A, Q2, _ = oracle.trace_descs(P, cell, ref)
        assert abs(res["E"][cell] - E) <= 1e-5 * abs(E)
        assert abs(res["A"][cell] - A) <= 1e-5 * abs(A)
        assert abs(res["Q"][cell] - (E - A)) <= 1e-5 * abs(E)
        rel.append(abs(res["Q"][cell] - (E - A)) / abs(E))
    return float(np.max(rel))


def test_grey_plain_parity(oracle):
    cfg = small_channel(R=64)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    T = workloads.field_grey32(3, tuple(cfg.n))
    S.set_field(0, workloads.ghosted(T))
    S.set_record(True)
    S.run()
    S.wait()
    seed = cfg.seed
    _check_rays_and_tallies(oracle, S, cfg, lambda P, c: oracle.gen_philox(P, seed, 0, c, 0, 64), 64)


def test_ck_plain_parity(oracle, ck_table_dir):
    tdir, t = ck_table_dir
    cfg = small_channel(R=48, spectral=0, Tw=573.0, table_dir=tdir)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    T = workloads.field_profile_noise(5, tuple(cfg.n))
    S.set_field(0, workloads.ghosted(T))
    S.set_record(True)
    S.run()
    S.wait()
    # derived tables against the oracle's own construction (bit-exact)
    Tmax = float(S.array("scalars")[0])
    assert Tmax == max(float(T.max()), 573.0)
    wq = workloads.gauss_legendre01(16)[1]
    parsed = read_reference_tables(tdir, gas=True)
    np.testing.assert_allclose(parsed["kq"], t["kq"], rtol=6e-7)   # writer round trip (%e: 7 digits)
    ref = oracle.build_proposal(0, parsed, Tmax, 48, wq)
    np.testing.assert_array_equal(S.array("Ib"), ref["Ib"])
    np.testing.assert_array_equal(S.array("kq"), ref["kq"])
    np.testing.assert_array_equal(S.array("inv_emax"), ref["inv_emax"])
    np.testing.assert_allclose(S.array("cdf_band"), ref["cdf_band"], rtol=0, atol=0)
    np.testing.assert_allclose(S.array("cdf_g"), ref["cdf_g"], rtol=0, atol=0)
    assert abs(S.Ibmax_d() - ref["Ibmax_d"]) <= 1e-12 * ref["Ibmax_d"]
    seed = cfg.seed
    _check_rays_and_tallies(oracle, S, cfg, lambda P, c: oracle.gen_philox(P, seed, 0, c, 0, 48), 48)


def _compat_particle_cfg(n=(16, 16, 16), photmax=96, grid_num=3):
    cfg = dnsmc_amd.default_config()
    d = dict(n=n, photmax=photmax, nvar=3, grid_num=grid_num, maxs=(5, 5, 100000, 100000, 100000, 100000),
             table_dir=particle_table_dir(), num_streams=16)
    return dnsmc_amd.config_from_dict(d, cfg)


def test_compat_particle_parity(oracle):
    """Reference sampling (XORWOW per (cell, block), Sobol directions, nearest
    node proposal) on the reference's own particle tables with 3 AMC levels."""
    cfg = _compat_particle_cfg()
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    x = S.array("centres0", 0)[1:-1].astype(np.float64)
    Tx = 500.0 - 2000.0 * x ** 2 + 2000.0 * x   # code/rad_main.f:68
    T = np.broadcast_to(Tx[:, None, None], tuple(cfg.n)).astype(np.float32)
    S.set_field(0, workloads.ghosted(np.ascontiguousarray(T)))
    S.set_record(True)
    S.run()
    S.wait()
    d = oracle_problem_dict(S, cfg)
    # reference proposal rows, Planck table and Ibmax against the oracle
    parsed = read_reference_tables(particle_table_dir(), gas=False)
    sc = S.array("scalars")
    ref = oracle.build_proposal(1, parsed, float(sc[0]), cfg.photmax, np.ones(1))
    np.testing.assert_array_equal(S.array("Ib"), ref["Ib"])
    np.testing.assert_array_equal(S.array("cdf_band"), ref["cdf_band"])
    np.testing.assert_array_equal(S.array("cdf_g"), ref["cdf_g"])
    assert float(sc[1]) == ref["kappamax"]
    assert abs(S.Ibmax_d() - ref["Ibmax_d"]) <= 1e-15 * ref["Ibmax_d"]
    # coarse levels from the GPU spline against the oracle's FP64 spline
    f0 = [S.array("centres%d" % a, 0) for a in range(3)]
    for l in range(1, cfg.grid_num):
        fc = [S.array("centres%d" % a, l) for a in range(3)]
        nc = [len(c) - 2 for c in fc]
        ref = oracle.interp3d(list(cfg.n), f0[0], f0[1], f0[2], d["levels"][0]["T"], nc, fc[0], fc[1], fc[2])
        np.testing.assert_allclose(d["levels"][l]["T"], ref, rtol=2e-7, atol=0)
    P = oracle.problem(d)
    R, nvar = cfg.photmax, cfg.nvar
    rays = S.rays(S.nemit * nvar * R)
    rec = S.records(S.nemit * nvar * R)
    n = tuple(cfg.n)
    cells = [0, 1, 17, 255, 1000, 2047, 4095]
    for cell in cells:
        i, j, k = cell % n[0], (cell // n[0]) % n[1], cell // (n[0] * n[1])
        for v in range(nvar):
            sub = ref_subsequence(i, j, k, v, n)
            ref = oracle.gen_compat(P, cfg.seed, sub, 1, v * R, cell, R)
            q0 = (cell * nvar + v) * R
            got = rays[q0:q0 + R]
            np.testing.assert_array_equal(got["w0"], ref["w0"])
            np.testing.assert_array_equal(got["w1"], ref["w1"])
            np.testing.assert_allclose(got["w"], ref["w"], rtol=2e-7)
            band, g = _decode(ref, d["nQ"])
            for r in range(0, R, 7):
                ut = (float(ref["w0"][r] & 0xFFFFFF) + 1.0) * 2.0 ** -24
                up = (float(ref["w1"][r] & 0xFFFFFF) + 1.0) * 2.0 ** -24
                o = oracle.trace_ray(P, i + 1, j + 1, k + 1, ut, up, int(band[r]), int(g[r]))
                assert rec["steps"][q0 + r] == o.steps
                assert rec["end"][q0 + r] == o.end_type
                assert rec["hash"][q0 + r] == o.hash


def test_isothermal_enclosure_zero():
    """Gas and walls at one temperature: Q == 0 to rounding (reciprocity)."""
    cfg = small_channel(R=64, Tw=1000.0)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    S.set_field(0, workloads.ghosted(np.full(tuple(cfg.n), 1000.0, np.float32)))
    S.run()
    r = S.results(0)
    assert np.all(np.abs(r["Q"]) <= 1e-5 * r["E"])


def test_grey_slab_closed_form(oracle):
    """Unperturbed isothermal grey slab: cell-centre OERM expectation is the
    plane-parallel exchange with the estimator's cut-off (E2 form without it)."""
    cfg = small_channel(n=(4, 32, 4), R=20000)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    S.set_field(0, workloads.ghosted(np.full(tuple(cfg.n), 1000.0, np.float32)))
    S.run()
    r = S.results(0)
    fy = S.array("faces1", 0).astype(np.float64)
    ny = cfg.n[1]
    kap = np.ones(ny)
    Ib = np.full(ny, 1000.0 ** 4)
    F = oracle.slab_exchange(fy, kap, Ib, 300.0 ** 4, 300.0 ** 4, tau_cut=-math.log(1e-3), npanel=4000)
    Qref = 4.0 * 5.670373e-08 * 1000.0 ** 4 * F
    Q = r["Q"].reshape(cfg.n[2], ny, cfg.n[0])
    z = (Q - Qref[None, :, None]) / r["sdev"].reshape(Q.shape)
    assert np.max(np.abs(z)) < 5.0
    assert abs(np.mean(z)) < 0.5


def test_grey_ref_tables_parity(oracle):
    """grey==1 (code/read_tables.cpp:104-183): 1000 synthesised bands of 15 1/cm,
    kappa = 1, temperature nodes spanning [450 K, 1.1 Tmax).  Rays, records and
    tallies against the oracle built from the library's synthesised arrays."""
    cfg = small_channel(R=32, spectral=1)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    T = workloads.field_grey32(4, tuple(cfg.n))
    S.set_field(0, workloads.ghosted(T))
    S.set_record(True)
    S.run()
    S.wait()
    sc = S.array("scalars")
    nB, nQ, nT = int(sc[3]), int(sc[4]), int(sc[5])
    assert (nB, nQ, nT) == (1000, 1, 55)
    Tmax = float(sc[0])
    # the synthesised node grid and Planck table (code/read_tables.cpp:107-125)
    Tnb = np.float32(450.0) + (np.float32(1.1) * np.float32(Tmax) - np.float32(450.0)) * np.arange(
        nT, dtype=np.float32) / np.float32(nT)
    nu = 15.0 * np.arange(nB) + 7.5
    ib = S.array("Ib").reshape(nB, nT).astype(np.float64)
    want = workloads.planck_band64(Tnb.astype(np.float64)[None, :], nu[:, None])
    np.testing.assert_allclose(ib, want, rtol=2e-5, atol=1e-30)
    np.testing.assert_array_equal(S.array("kq"), np.ones(nB * nT, np.float32))
    seed = cfg.seed
    _check_rays_and_tallies(oracle, S, cfg, lambda P, c: oracle.gen_philox(P, seed, 0, c, 0, 32), 32)


def test_grey_ref_tables_slab_closed_form(oracle):
    """grey==1 on the isothermal slab: every band has kappa = 1, so the
    wall-arrival fractions are the grey slab's (cut-off included) and
    Q(y) = 4 pi sum_b dnu_b (Ib_b(T) - Ib_b(Tw)) F(y); the emission
    4 pi sum_b dnu_b Ib_b(T) is 4 sigma T^4 to the table's interpolation."""
    cfg = small_channel(n=(4, 32, 4), R=20000, spectral=1)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    S.set_field(0, workloads.ghosted(np.full(tuple(cfg.n), 1000.0, np.float32)))
    S.run()
    r = S.results(0)
    sc = S.array("scalars")
    nB, nT = int(sc[3]), int(sc[5])
    inv_dT, mT0 = np.float32(sc[6]), np.float32(sc[7])
    pos = float(np.float32(np.float32(1000.0) * inv_dT + mT0))
    t0 = min(int(pos), nT - 2)
    fr = pos - t0
    ib = S.array("Ib").reshape(nB, nT).astype(np.float64)
    ib_gas = ib[:, t0] + fr * (ib[:, t0 + 1] - ib[:, t0])
    ib_w = S.array("Ibw")[2 * nB:3 * nB].astype(np.float64)   # the y = 0 wall (axis 1, low side) at 300 K
    np.testing.assert_allclose(ib_w, workloads.planck_band64(300.0, 15.0 * np.arange(nB) + 7.5), rtol=2e-5,
                               atol=1e-30)
    dnu = 1500.0
    emis = 4.0 * math.pi * dnu * ib_gas.sum()
    np.testing.assert_allclose(r["E"], emis, rtol=2e-5)
    assert emis == pytest.approx(4.0 * 5.670373e-08 * 1000.0 ** 4, rel=1e-3)
    fy = S.array("faces1", 0).astype(np.float64)
    ny = cfg.n[1]
    F = oracle.slab_exchange(fy, np.ones(ny), np.ones(ny), 0.0, 0.0, tau_cut=-math.log(1e-3), npanel=4000)
    Qref = 4.0 * math.pi * dnu * (ib_gas - ib_w).sum() * F
    Q = r["Q"].reshape(cfg.n[2], ny, cfg.n[0])
    z = (Q - Qref[None, :, None]) / r["sdev"].reshape(Q.shape)
    assert np.max(np.abs(z)) < 5.0
    assert abs(np.mean(z)) < 0.5


def test_legacy_abi_fortran_layout(tmp_path):
    """mc_gpu_/get_results_ through the Fortran-layout boundary."""
    conf = tmp_path / "rad.conf"
    conf.write_text(legacy_conf_text())
    os.environ["DNSMC_CONFIG"] = str(conf)
    cfg = dnsmc_amd.parse_config(conf.read_text())
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    x = S.array("centres0", 0)
    Tg = np.empty((18, 18, 18), np.float32, order="F")
    Tg[...] = (500.0 - 2000.0 * x ** 2 + 2000.0 * x)[:, None, None]
    S.set_field(0, Tg)
    S.run()
    Qs, Vs = S.results_fortran(0)
    Q = np.full((18, 18, 18), -7.0, np.float32, order="F")
    V = np.full((18, 18, 18), -7.0, np.float32, order="F")
    dnsmc_amd.mc_gpu(Tg, 1)
    dnsmc_amd.get_results(Q, V)
    np.testing.assert_array_equal(Q[1:-1, 1:-1, 1:-1], Qs[1:-1, 1:-1, 1:-1])
    np.testing.assert_array_equal(V[1:-1, 1:-1, 1:-1], Vs[1:-1, 1:-1, 1:-1])
    assert np.all(Q[0] == -7.0) and np.all(Q[:, -1] == -7.0)


@pytest.mark.parametrize("tally", [1, 2, 3])
def test_tally_modes_agree(tally):
    cfg0 = small_channel(R=200)
    cfgt = small_channel(R=200, tally=tally)
    T = workloads.ghosted(workloads.field_grey32(4, tuple(cfg0.n)))
    out = []
    for c in (cfg0, cfgt):
        S = dnsmc_amd.Solver(c)
        S.load_tables()
        S.set_field(0, T)
        S.run()
        out.append(S.results(0))
    tol = {1: 1e-12, 2: 2e-6, 3: 1e-5}[tally]
    np.testing.assert_allclose(out[1]["E"], out[0]["E"], rtol=tol)
    np.testing.assert_allclose(out[1]["A"], out[0]["A"], rtol=tol)


@pytest.mark.parametrize("R", [2, 3, 8, 24])
def test_tally_ring_with_few_rays_per_cell(ck_table_dir, R):
    """The per-wave tally ring (atomic tallies) against the deterministic
    per-ray FP64 sums when a 64-ray queue batch spans many processing slots
    (R = 2: 32 slots, more than the ring holds, so entries are evicted while
    their rays are in flight and those rays tally straight to the slot): E,
    A and the variance sums agree to FP64 rounding, on the fine grid and with
    coarsening levels."""
    tdir, _ = ck_table_dir
    for amc in ({}, dict(grid_num=3, maxs=(3, 3, 100000))):
        out = []
        for tally in (0, 1):
            cfg = small_channel(n=(12, 16, 8), R=R, spectral=0, Tw=573.0, table_dir=tdir, tally=tally, **amc)
            S = dnsmc_amd.Solver(cfg)
            S.load_tables()
            S.set_field(0, workloads.ghosted(workloads.field_profile_noise(7, tuple(cfg.n))))
            S.run()
            out.append((S.results(0), S.stats()))
            S.close()
        (r0, s0), (r1, s1) = out
        assert s0["n_steps"] == s1["n_steps"]
        np.testing.assert_allclose(r0["E"], r1["E"], rtol=1e-12, atol=0)
        np.testing.assert_allclose(r0["A"], r1["A"], rtol=1e-12, atol=0)
        np.testing.assert_allclose(r0["sdev"], r1["sdev"], rtol=1e-9, atol=1e-300)


@pytest.mark.parametrize("photmax", [24, 96])
def test_tally_ring_compat_mode(photmax):
    """The reference sampling (several tally slots per emitting cell, one per
    variance block): ring tallies against the deterministic FP64 sums on the
    particle tables with coarsening, E, A and the block variances."""
    out = []
    for tally in (0, 1):
        cfg = _compat_particle_cfg(n=(16, 12, 12), photmax=photmax, grid_num=3)
        cfg.tally = tally
        S = dnsmc_amd.Solver(cfg)
        S.load_tables()
        x = S.array("centres0", 0)
        Tg = np.empty((18, 14, 14), np.float32, order="F")
        Tg[...] = (500.0 - 2000.0 * x ** 2 + 2000.0 * x)[:, None, None]
        S.set_field(0, Tg)
        S.run()
        out.append((S.results(0), S.stats()))
        S.close()
    (r0, s0), (r1, s1) = out
    assert s0["n_steps"] == s1["n_steps"]
    np.testing.assert_allclose(r0["E"], r1["E"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(r0["A"], r1["A"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(r0["sdev"], r1["sdev"], rtol=1e-9, atol=1e-300)


def test_table_placement_bit_identical(ck_table_dir, monkeypatch):
    """The kappa/Ib slice in LDS and the global-memory tables give the same
    per-ray values, so the deterministic FP64 tallies agree exactly."""
    tdir, _ = ck_table_dir
    cfg = small_channel(n=(12, 16, 10), R=96, spectral=0, Tw=573.0, table_dir=tdir, tally=1)
    T = workloads.ghosted(workloads.field_profile_noise(8, tuple(cfg.n)))
    out, modes = [], []
    for force in (None, "global"):
        if force:
            monkeypatch.setenv("DNSMC_TABLES", force)
        S = dnsmc_amd.Solver(cfg)
        S.load_tables()
        S.set_field(0, T)
        S.run()
        out.append(S.results(0))
        modes.append(S.stats()["table_mode"])
    assert modes[0] == 2 and modes[1] == 0
    np.testing.assert_array_equal(out[0]["E"], out[1]["E"])
    np.testing.assert_array_equal(out[0]["A"], out[1]["A"])


@pytest.mark.parametrize("sobol", [1, 0])
def test_compat_sorted_gas_parity(oracle, ck_table_dir, sobol):
    """Sorted compat mode (srt == 1) on the correlated-k table: every bundle
    descriptor bit-exact against the oracle's kernel_find / wave_find replay,
    per-ray traversal integers on a sample, and the block-averaged tallies."""
    tdir, _ = ck_table_dir
    n = (16, 8, 6)
    R, nvar = 400, 3
    cfg = workloads.channel_config(n, spectral=0, Tw=573.0, table_dir=tdir, mode=1, srt=1, sobol=sobol,
                                   photmax=R, nvar=nvar, num_streams=16)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    S.set_field(0, workloads.ghosted(workloads.field_profile_noise(7, n)))
    S.set_record(True)
    S.run()
    S.wait()
    d = oracle_problem_dict(S, cfg)
    P = oracle.problem(d)
    order = oracle.sort_idx(read_reference_tables(tdir, gas=True)["kavg"])
    rays = S.rays(S.nemit * nvar * R)
    rec = S.records(S.nemit * nvar * R)
    res = S.results(0)
    for cell in [0, 1, 37, 200, 511, 767]:
        i, j, k = cell % n[0], (cell // n[0]) % n[1], cell // (n[0] * n[1])
        dq = []
        for v in range(nvar):
            ref = oracle.gen_compat_sorted(P, cfg.seed, ref_subsequence(i, j, k, v, n), sobol, v * R, cell, R, order)
            q0 = (cell * nvar + v) * R
            got = rays[q0:q0 + R]
            np.testing.assert_array_equal(got["w0"], ref["w0"])
            np.testing.assert_array_equal(got["w1"], ref["w1"])
            np.testing.assert_allclose(got["w"], ref["w"], rtol=2e-7)
            band, g = _decode(ref, d["nQ"])
            for r in range(0, R, 23):
                ut = (float(ref["w0"][r] & 0xFFFFFF) + d["u_off"]) * 2.0 ** -24
                up = (float(ref["w1"][r] & 0xFFFFFF) + d["u_off"]) * 2.0 ** -24
                o = oracle.trace_ray(P, i + 1, j + 1, k + 1, ut, up, int(band[r]), int(g[r]))
                assert rec["steps"][q0 + r] == o.steps
                assert rec["end"][q0 + r] == o.end_type
                assert rec["hash"][q0 + r] == o.hash
            E, A, _, _ = oracle.trace_descs(P, cell, ref)
            dq.append((E, A))
        E = np.mean([e for e, _ in dq])
        A = np.mean([a for _, a in dq])
        assert abs(res["E"][cell] - E) <= 1e-5 * abs(E)
        assert abs(res["Q"][cell] - (E - A)) <= 1e-5 * abs(E)


@pytest.mark.parametrize("spectral", [0, 3])
def test_fp64_traversal_matches_fp64_oracle(oracle, ck_table_dir, spectral):
    """precision 1: every ray in FP64 (directions, lookups, transport, weights)
    against the oracle's FP64 estimator on the same Philox rays.  Per-ray
    results can differ only through the last bit of cos/sin/exp (device math
    library vs glibc), so tallies agree to 1e-12 relative; step counts and end
    types are equal ray by ray."""
    tdir, _ = ck_table_dir
    n, R = (8, 12, 8), 32
    cfg = workloads.channel_config(n, rays_per_cell=R, spectral=spectral, Tw=573.0 if spectral == 0 else 300.0,
                                   table_dir=tdir, precision=1, tally=1)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    T = workloads.field_profile_noise(5, n) if spectral == 0 else workloads.field_grey32(5, n)
    S.set_field(0, workloads.ghosted(T))
    S.set_record(True)
    S.run()
    st = S.stats()
    res = S.results(0)
    rec = S.records(S.nemit * R)
    d = oracle_problem_dict(S, cfg)
    P = oracle.problem(d)
    E, A, Q2, ns = oracle.mc_philox(P, cfg.seed, 0, R, f64=True)
    np.testing.assert_allclose(res["E"], E, rtol=1e-12, atol=0)
    np.testing.assert_allclose(res["A"], A, rtol=1e-12, atol=0)
    assert np.all(np.abs(res["Q"] - (E - A)) <= 1e-12 * np.abs(E))
    assert st["n_steps"] == ns
    assert int(np.sum(rec["steps"])) == ns
    # FP32 traversal on the same rays: the precision error of the FP32 path
    cfg32 = workloads.channel_config(n, rays_per_cell=R, spectral=spectral, Tw=cfg.Tw[1][0], table_dir=tdir,
                                     precision=0, tally=1)
    S32 = dnsmc_amd.Solver(cfg32)
    S32.load_tables()
    S32.set_field(0, workloads.ghosted(T))
    S32.run()
    r32 = S32.results(0)
    assert np.max(np.abs(r32["E"] - E) / np.abs(E)) < 1e-5
    assert np.max(np.abs(r32["Q"] - (E - A)) / np.abs(E)) < 1e-5


@pytest.mark.parametrize("spectral,grid_num,part", [(0, 1, None), (0, 2, None), (3, 1, None), (0, 2, 37)])
def test_fused_generation_matches_batched(ck_table_dir, monkeypatch, spectral, grid_num, part):
    """The plain estimator's rays drawn inside the traversal (one launch for
    all snapshots) against descriptors drawn by raygen_plain (batched
    launches, which the oracle pins ray by ray): same rays, so equal step
    totals and tallies equal up to the order of the FP64 atomic additions.
    part: slots per fused launch (the split a snapshot of more than 2^32
    rays takes), here 37 of the 4096 slots per snapshot."""
    tdir, _ = ck_table_dir
    if part is not None:
        monkeypatch.setenv("DNSMC_FUSED_SLOTS", str(part))
    n, R, ns = (16, 24, 16), 48, 3
    maxs = (6, 100000, 100000, 100000, 100000, 100000)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("DNSMC_FUSED", fused)
        cfg = workloads.channel_config(n, rays_per_cell=R, spectral=spectral, Tw=573.0 if spectral == 0 else 300.0,
                                       table_dir=tdir, nsnap=ns, grid_num=grid_num, maxs=maxs,
                                       batch_rays=R * 1000)
        S = dnsmc_amd.Solver(cfg)
        S.load_tables()
        for s in range(ns):
            T = workloads.field_fourier(60 + s, n) if spectral == 0 else workloads.field_grey32(60 + s, n)
            S.set_field(s, workloads.ghosted(T))
        S.run()
        st = S.stats()
        out[fused] = (st, [S.results(s) for s in range(ns)])
        S.close()
    (sf, rf), (sb, rb) = out["1"], out["0"]
    assert sf["n_trace_launches"] == (1 if part is None else -(-n[0] * n[1] * n[2] // part))
    assert sb["n_trace_launches"] > ns
    assert sf["n_steps"] == sb["n_steps"]
    for s in range(ns):
        np.testing.assert_allclose(rf[s]["E"], rb[s]["E"], rtol=1e-12, atol=0)
        np.testing.assert_allclose(rf[s]["A"], rb[s]["A"], rtol=1e-12, atol=0)
        assert np.all(np.abs(rf[s]["Q"] - rb[s]["Q"]) <= 1e-12 * np.abs(rb[s]["E"]))


@pytest.mark.parametrize("grid_num", [1, 2])
def test_long_rays_first_changes_only_the_order(ck_table_dir, monkeypatch, grid_num):
    """Long rays first: a fused launch lists the rays it would start last (the
    final half of each of the 8 XCD ranges of its last snapshot, with several
    snapshots per launch) whose
    polar direction cosine is below 0.02 in magnitude, traces those first and
    skips them where their super-chunks come up.  Against DNSMC_LONG_FIRST=0:
    every ray traced once (equal step totals), tallies equal up to the order
    of the FP64 atomic additions, and the listed rays about 2 % of that final
    half (binomial count, 5 sigma)."""
    tdir, _ = ck_table_dir
    n, R, ns = (16, 24, 16), 48, 2
    maxs = (6, 100000, 100000, 100000, 100000, 100000)
    out = {}
    for lf in ("0", "1"):
        monkeypatch.setenv("DNSMC_LONG_FIRST", lf)
        cfg = workloads.channel_config(n, rays_per_cell=R, spectral=0, Tw=573.0, table_dir=tdir, nsnap=ns,
                                       grid_num=grid_num, maxs=maxs)
        S = dnsmc_amd.Solver(cfg)
        S.load_tables()
        for s in range(ns):
            S.set_field(s, workloads.ghosted(workloads.field_fourier(80 + s, n)))
        S.run()
        st = S.stats()
        out[lf] = (st, [S.results(s) for s in range(ns)])
        S.close()
    (s0, r0), (s1, r1) = out["0"], out["1"]
    assert s0["n_listed"] == 0
    # the final half of the 8 XCD ranges of 512-ray super-chunks
    nrays, sc = n[0] * n[1] * n[2] * R, 512
    per_x = -(-(-(-nrays // sc)) // 8)
    clo, D = per_x - (-(-per_x // 2)), per_x * sc
    fin = sum(max(0, min((x + 1) * D, nrays) - (x * D + clo * sc)) for x in range(8))
    p = 2 * 167772 / 2.0 ** 24
    assert abs(s1["n_listed"] - fin * p) < 5 * math.sqrt(fin * p * (1 - p)), (s1["n_listed"], fin * p)
    assert s1["n_steps"] == s0["n_steps"]
    for s in range(ns):
        np.testing.assert_allclose(r1[s]["E"], r0[s]["E"], rtol=1e-12, atol=0)
        np.testing.assert_allclose(r1[s]["A"], r0[s]["A"], rtol=1e-12, atol=0)


def test_long_ray_list_overflow_falls_back_to_plain_order(ck_table_dir, monkeypatch):
    """Lists of 8 rays (DNSMC_LONG_CAP=8) against the ~370 that each XCD
    range's half finds: the launch turns the listing off (nothing listed,
    nothing skipped) and traces every ray in plain claim order, so the step
    totals and tallies equal those of DNSMC_LONG_FIRST=0."""
    tdir, _ = ck_table_dir
    n, R, ns = (16, 24, 16), 48, 2
    out = {}
    for key, env in (("off", {"DNSMC_LONG_FIRST": "0"}), ("overflow", {"DNSMC_LONG_CAP": "8"})):
        for k in ("DNSMC_LONG_FIRST", "DNSMC_LONG_CAP"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        cfg = workloads.channel_config(n, rays_per_cell=R, spectral=0, Tw=573.0, table_dir=tdir, nsnap=ns)
        S = dnsmc_amd.Solver(cfg)
        S.load_tables()
        for s in range(ns):
            S.set_field(s, workloads.ghosted(workloads.field_fourier(90 + s, n)))
        S.run()
        st = S.stats()
        out[key] = (st, [S.results(s) for s in range(ns)])
        S.close()
    (s0, r0), (s1, r1) = out["off"], out["overflow"]
    assert s0["n_listed"] == 0 and s1["n_listed"] == 0
    assert s1["n_steps"] == s0["n_steps"]
    for s in range(ns):
        np.testing.assert_allclose(r1[s]["E"], r0[s]["E"], rtol=1e-12, atol=0)
        np.testing.assert_allclose(r1[s]["A"], r0[s]["A"], rtol=1e-12, atol=0)


def test_fp64_traversal_with_coarsening_matches_fp64_oracle(oracle, ck_table_dir):
    """precision 1 through two coarsening levels (5 fine steps, then the
    coarse spline field): tallies against the oracle's FP64 estimator."""
    tdir, _ = ck_table_dir
    n, R = (16, 24, 16), 24
    cfg = workloads.channel_config(n, rays_per_cell=R, spectral=0, Tw=573.0, table_dir=tdir, precision=1, tally=1,
                                   grid_num=2, maxs=(5, 100000, 100000, 100000, 100000, 100000))
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    S.set_field(0, workloads.ghosted(workloads.field_fourier(9, n)))
    S.run()
    st = S.stats()
    res = S.results(0)
    P = oracle.problem(oracle_problem_dict(S, cfg))
    E, A, Q2, ns = oracle.mc_philox(P, cfg.seed, 0, R, f64=True)
    np.testing.assert_allclose(res["E"], E, rtol=1e-12, atol=0)
    np.testing.assert_allclose(res["A"], A, rtol=1e-12, atol=0)
    assert st["n_steps"] == ns


def test_reference_driver_resfortm():
    """The reference's only checked-in result of the hot path, code/resfortm.txt:
    its standalone driver (code/rad_main.f) with the shipped configuration
    (code/param.h: 64^3, particle tables, 5 coarsening levels, photmax 20000,
    nVar 3, Sobol directions; T = 500 - 2000 x^2 + 2000 x, code/rad_main.f:68),
    plane means over (j, k) of Q, of its standard error and the spatial RMS.
    Our compat mode uses its own XORWOW seeding, so the agreement is
    statistical: each plane mean within 3 per-cell standard errors (the Sobol
    directions are shared by all cells, so plane means are not averaged down),
    and the error estimates themselves within 10 % / 15 %."""
    from helpers import read_resfortm
    cfg = dnsmc_amd.config_from_dict(dict(table_dir=particle_table_dir()), dnsmc_amd.default_config())
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    x = S.array("centres0", 0)[1:-1].astype(np.float64)
    ref = read_resfortm()
    np.testing.assert_allclose(x, ref[:, 0], rtol=0, atol=8e-8)
    Tx = 500.0 - 2000.0 * x ** 2 + 2000.0 * x
    T = np.broadcast_to(Tx[:, None, None], tuple(cfg.n)).astype(np.float32)
    S.set_field(0, workloads.ghosted(np.ascontiguousarray(T)))
    S.run()
    r = S.results(0)
    n = tuple(cfg.n)
    Q = r["Q"].reshape(n[2], n[1], n[0])
    Qm = Q.mean(axis=(0, 1))
    Vm = r["sdev"].reshape(n[2], n[1], n[0]).mean(axis=(0, 1))
    rms = Q.std(axis=(0, 1))
    assert np.all(np.abs(Qm - ref[:, 1]) <= 3.0 * ref[:, 2]), np.max(np.abs(Qm - ref[:, 1]) / ref[:, 2])
    assert np.mean(np.abs(Qm[2:-2] / ref[2:-2, 1] - 1.0)) < 5e-4
    np.testing.assert_allclose(Vm, ref[:, 2], rtol=0.10)
    np.testing.assert_allclose(rms, ref[:, 3], rtol=0.15)
