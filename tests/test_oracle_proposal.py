"""CPU pins of the oracle's plain correlated-k proposal (orc_build_proposal
mode 0), which the plain estimator and the throughput workload run.

Reference semantics: the bundle weight Ibmax * kappa_i Ib_i / (kappa_max Ib_max)
(code/mc_gpu.cu:63, :163-169) and kappamax interpolated at Tmax
(code/read_tables.cpp:34-36).  The plain proposal draws (band, g) from a CDF
built at Tmax itself, so every bundle's weight is w(bg; T_i) =
Ibmax kappa(T_i,b,g) Ib(T_i,b) / (kappa(Tmax,b,g) Ib(Tmax,b)) with
Ibmax = (total emission at Tmax) / R.

Three checks, none of which re-types the oracle's arithmetic:

1. the emission identity R sum_bg p(bg) w(bg; T_i) =
   4 pi sum_b dnu_b sum_g w_g kappa(T_i,b,g) Ib(T_i,b), the right side summed
   in numpy from the parsed table files with the Planck closed form, at several
   Tmax and T_i, on the 100x16x69 table, on a variant with unequal band widths
   and with asymmetric g weights;
2. its grey limit: kappa = 1 on 1000 bands of 15 1/cm (the reference's grey==1
   band layout, code/read_tables.cpp:104-183) gives 4 kappa sigma T^4;
3. the optically thin closed form Q_i = 4 pi sum_b dnu_b sum_g w_g kappa_i
   (Ib_b(T_i) - Ib_b(T_w)) through the full estimator (orc_mc_philox), z < 4.

Each check is shown to fail when a mistake is planted in the oracle (the width
of the next band, the g weights reversed, the g CDF shifted by one entry).
"""
import math

import numpy as np
import pytest

from helpers import read_reference_tables

SIGMA = 5.670373e-08
# Planck constants of the reference's I_blackC (code/memory_copy.cu:294-295)
C1 = 3.741771790075259e-16
C2 = 0.014387741858429


def planck64(T, nu):
    x = np.asarray(nu, np.float64) * 100.0
    return C1 / math.pi * x ** 3 / np.expm1(C2 * x / np.asarray(T, np.float64))


def lerp_nodes(Tnb, rows, T):
    """Linear interpolation in T along the last axis (the tables' node grid)."""
    Tnb = np.asarray(Tnb, np.float64)
    nT = len(Tnb)
    pos = (T - Tnb[0]) / (Tnb[1] - Tnb[0])
    pos = min(max(pos, 0.0), nT - 1.0)
    t0 = min(int(pos), nT - 2)
    fr = pos - t0
    rows = np.asarray(rows, np.float64)
    return rows[..., t0] + fr * (rows[..., t0 + 1] - rows[..., t0])


def emission_numpy(t, wq, T):
    """4 pi sum_b dnu_b sum_g w_g kappa(T,b,g) Ib(T,b), straight from the table
    data: kappa rows lerped in T, Planck at the band centre on the T nodes."""
    kq = np.asarray(t["kq"], np.float64)                     # [b][t][g]
    Ib_nodes = planck64(np.asarray(t["Tnb"], np.float64)[None, :], np.asarray(t["wvc"], np.float64)[:, None])
    ib = lerp_nodes(t["Tnb"], Ib_nodes, T)                   # [b]
    kap = lerp_nodes(t["Tnb"], np.transpose(kq, (0, 2, 1)), T)   # [b][g]
    dnu = (np.asarray(t["wvr"], np.float64) - np.asarray(t["wvl"], np.float64)) * 100.0
    return 4.0 * math.pi * float(np.sum(dnu * ib * (kap @ np.asarray(wq, np.float64))))


def proposal_emission(pr, nB, nQ, nT, Tnb, R, T):
    """R sum_bg p(bg) w(bg; T): p from the oracle's band and g CDFs, w as the
    kernels weigh a bundle (Ibmax * kappa_i Ib_i * 1/(kappa_max Ib_max)), with
    kappa_i and Ib_i lerped from the oracle's own table arrays."""
    cb = np.asarray(pr["cdf_band"], np.float64)
    cg = np.asarray(pr["cdf_g"], np.float64).reshape(nB, nQ)
    pb = np.diff(np.concatenate([[0.0], cb]))
    pg = np.diff(np.concatenate([np.zeros((nB, 1)), cg], axis=1), axis=1)
    p = pb[:, None] * pg
    kap = lerp_nodes(Tnb, np.asarray(pr["kq"], np.float64).reshape(nB, nQ, nT), T)
    ib = lerp_nodes(Tnb, np.asarray(pr["Ib"], np.float64).reshape(nB, nT), T)
    w = pr["Ibmax_d"] * kap * ib[:, None] * np.asarray(pr["inv_emax"], np.float64).reshape(nB, nQ)
    return R * float(np.sum(p * w))


def uneven_table(t, seed=5):
    """The same table with band widths varying by a factor of three (band edges
    contiguous, centres mid-band) and a temperature dependence of kappa that
    varies with g: kappa * (T / 1000 K)^(0.6 (gq_g - 1/2)).  In the synthetic
    table kappa(T_i)/kappa(Tmax) depends on the band only, which would leave
    any mistake inside a band's g distribution invisible to the emission
    identity.  The plain proposal reads only Tnb, the band edges and centres
    and kq, so the other columns are left as they are."""
    u = dict(t)
    gq = np.asarray(t["gq"], np.float64)
    u["kq"] = t["kq"] * (t["Tnb"][None, :, None] / 1000.0) ** (0.6 * (gq[None, None, :] - 0.5))
    nB = len(t["wvl"])
    rng = np.random.default_rng(seed)
    w = 0.5 + rng.random(nB)
    w *= (t["wvr"][-1] - t["wvl"][0]) / w.sum()
    edges = t["wvl"][0] + np.concatenate([[0.0], np.cumsum(w)])
    u["wvl"] = edges[:-1].copy()
    u["wvr"] = edges[1:].copy()
    u["wvc"] = 0.5 * (edges[:-1] + edges[1:])
    return u


@pytest.fixture(scope="module")
def tables(ck_table_dir, tmp_path_factory):
    from dnsmc_amd import workloads
    d, t = ck_table_dir
    even = read_reference_tables(d)
    d2 = str(tmp_path_factory.mktemp("ck_uneven"))
    workloads.write_reference_tables(uneven_table(t), d2)
    uneven = read_reference_tables(d2)
    gq, wq = workloads.gauss_legendre01(16)
    rng = np.random.default_rng(9)
    wq_asym = rng.random(16) + 0.2
    wq_asym /= wq_asym.sum()
    return dict(even=even, uneven=uneven, wq=wq, wq_asym=wq_asym)


def _identity_errors(oracle, t, wq, R=100):
    nB, nT, nQ = t["kq"].shape
    errs = []
    for Tmax in (955.0, 1000.0, 1237.3, 1812.5):
        pr = oracle.build_proposal(0, t, Tmax, R, wq)
        for frac in (1.0, 0.93, 0.8, 0.61, 0.45):
            Ti = float(np.float32(Tmax * frac))
            if Ti < t["Tnb"][0]:
                continue
            lhs = proposal_emission(pr, nB, nQ, nT, t["Tnb"], R, Ti)
            rhs = emission_numpy(t, wq, Ti)
            errs.append(abs(lhs / rhs - 1.0))
    return np.array(errs)


def test_plain_proposal_emission_identity(oracle, tables):
    for name, wkey in (("even", "wq"), ("uneven", "wq"), ("uneven", "wq_asym"), ("even", "wq_asym")):
        e = _identity_errors(oracle, tables[name], tables[wkey])
        assert len(e) >= 18
        assert e.max() < 2e-6, (name, wkey, e.max())


def test_plain_proposal_cdf_rows(oracle, tables):
    t = tables["uneven"]
    nB, nT, nQ = t["kq"].shape
    pr = oracle.build_proposal(0, t, 1111.0, 64, tables["wq"])
    cb = np.asarray(pr["cdf_band"])
    cg = np.asarray(pr["cdf_g"]).reshape(nB, nQ)
    assert np.all(np.diff(cb) >= 0) and cb[-1] == 1.0 and cb[0] > 0
    assert np.all(np.diff(cg, axis=1) >= 0) and np.all(cg[:, -1] == 1.0)
    # the band probabilities are the bands' shares of the emission at Tmax
    share = np.array([emission_numpy(dict(t, kq=t["kq"][b:b + 1], wvl=t["wvl"][b:b + 1], wvr=t["wvr"][b:b + 1],
                                          wvc=t["wvc"][b:b + 1]), tables["wq"], 1111.0) for b in range(nB)])
    np.testing.assert_allclose(np.diff(np.concatenate([[0.0], cb])), share / share.sum(), rtol=2e-4, atol=2e-7)


@pytest.mark.parametrize("mistake,table,wkey", [(1, "uneven", "wq"), (2, "even", "wq_asym"), (3, "uneven", "wq")])
def test_identity_fails_on_planted_mistakes(oracle, tables, mistake, table, wkey):
    oracle.plant_mistake(mistake)
    try:
        e = _identity_errors(oracle, tables[table], tables[wkey])
    finally:
        oracle.plant_mistake(0)
    assert e.max() > 1e-2, (mistake, e.max())


def planck_tail(x):
    """Fraction of sigma T^4 above x = C2 nu / T: (15/pi^4) sum_n e^(-n x)/n
    (x^3 + 3x^2/n + 6x/n^2 + 6/n^3), the textbook series."""
    return 15.0 / math.pi ** 4 * sum(math.exp(-n * x) / n * (x ** 3 + 3 * x * x / n + 6 * x / n ** 2 + 6 / n ** 3)
                                     for n in range(1, 200))


def test_grey_limit_is_four_kappa_sigma_T4(oracle):
    """kappa = 1 on the reference's grey==1 band layout (1000 bands of 15 1/cm
    up to 15000 1/cm, code/read_tables.cpp:104-183): the tabulated emission
    4 pi sum_b dnu_b Ib_b(T) must be 4 sigma T^4 less the Planck tail above
    15000 1/cm (3e-4 at 1500 K), through the proposal at several Tmax."""
    nB, nT = 1000, 69
    Tnb = np.arange(300.0, 2000.0 + 1.0, 25.0)
    wvl = 15.0 * np.arange(nB)
    t = dict(Tnb=Tnb, kP=np.ones(nT), wvl=wvl, wvr=wvl + 15.0, wvc=wvl + 7.5, kq=np.ones((nB, nT, 1)),
             prob=np.zeros((nT, nB)), probg=np.ones((nT, nB, 1)))
    R = 100
    # the temperatures are table nodes, where Ib is the Planck value itself
    for Tmax, Ti in ((1500.0, 1500.0), (1500.0, 1000.0), (1237.3, 600.0), (1812.5, 1200.0)):
        pr = oracle.build_proposal(0, t, Tmax, R, np.ones(1))
        got = proposal_emission(pr, nB, 1, nT, Tnb, R, Ti)
        want = 4.0 * SIGMA * Ti ** 4 * (1.0 - planck_tail(C2 * 1.5e6 / Ti))
        assert got == pytest.approx(want, rel=3e-5), (Tmax, Ti, got / want - 1.0)


# ----------------------------------------------------- optically thin ----
def thin_problem(oracle, t, wq, T, Tw, R, n=(3, 16, 3), L=(1.0, 2.0, 1.0), scale=1e-8):
    """Channel with black walls at Tw (y), periodic x and z, kappa scaled to the
    optically thin limit; the plain proposal built at Tmax = max(T, Tw)."""
    ts = dict(t, kq=np.asarray(t["kq"], np.float64) * scale)
    fx, cx = oracle.grid("uniform", n[0], L[0])
    fy, cy = oracle.grid("tanh", n[1], L[1], beta=1.5)
    fz, cz = oracle.grid("uniform", n[2], L[2])
    Tmax = float(max(float(np.max(T)), Tw))
    pr = oracle.build_proposal(0, ts, Tmax, R, wq)
    nB, nT, nQ = ts["kq"].shape
    Tnb = np.asarray(ts["Tnb"], np.float32)
    inv_dT = np.float32(1.0) / np.float32(Tnb[1] - Tnb[0])
    mT0 = -Tnb[0] * inv_dT
    ibw = np.array([oracle.planck_band(Tw, float(nu)) for nu in np.asarray(ts["wvc"], np.float32)], np.float32)
    d = dict(
        levels=[dict(n=list(n), faces=[fx, fy, fz], T=np.asarray(T, np.float32), maxs=100000, centres=[cx, cy, cz])],
        bc=[[1, 1], [2, 2], [1, 1]], polar_axis=1, tau_cut=float(np.float32(-math.log(1e-3))),
        spectral=0, nB=nB, nQ=nQ, nT=nT, inv_dT=float(inv_dT), mT0=float(mT0),
        kq=pr["kq"], Ib=pr["Ib"], Ibw=np.tile(ibw, 6), kappa_grey=0.0,
        cdf_band=pr["cdf_band"], cdf_g=pr["cdf_g"], inv_emax=pr["inv_emax"],
        Ibmax=float(np.float32(pr["Ibmax_d"])), Ibmax_d=pr["Ibmax_d"], Tmax=Tmax, g_search=1, u_off=0.5,
    )
    return d, ts


def thin_closed_form(t, wq, T, Tw):
    """Q_i = 4 pi sum_b dnu_b sum_g w_g kappa(T_i,b,g) (Ib_b(T_i) - Ib_b(T_w)),
    wall intensities exact at Tw (the walls' emission is not interpolated,
    code/mc_gpu.cu:541-561)."""
    kq = np.asarray(t["kq"], np.float64)
    Ib_nodes = planck64(np.asarray(t["Tnb"], np.float64)[None, :], np.asarray(t["wvc"], np.float64)[:, None])
    ibw = planck64(Tw, np.asarray(t["wvc"], np.float64))
    dnu = (np.asarray(t["wvr"], np.float64) - np.asarray(t["wvl"], np.float64)) * 100.0
    out = np.zeros(len(T))
    for i, Ti in enumerate(np.asarray(T, np.float64)):
        ib = lerp_nodes(t["Tnb"], Ib_nodes, Ti)
        kap = lerp_nodes(t["Tnb"], np.transpose(kq, (0, 2, 1)), Ti)
        out[i] = 4.0 * math.pi * float(np.sum(dnu * (ib - ibw) * (kap @ np.asarray(wq, np.float64))))
    return out


def _thin_z(oracle, t, wq, R=20000, seed=31):
    n = (3, 16, 3)
    rng = np.random.default_rng(seed)
    T = np.float32(700.0 + 500.0 * rng.random(n[0] * n[1] * n[2]))
    Tw = 573.0
    d, ts = thin_problem(oracle, t, wq, T, Tw, R, n=n)
    P = oracle.problem(d)
    E, A, Q2, _ = oracle.mc_philox(P, seed, 0, R)
    Q = E - A
    var = (R * Q2 - Q ** 2) / (R - 1)
    want = thin_closed_form(ts, wq, T.astype(np.float64), Tw)
    z = (Q - want) / np.sqrt(var)
    ztot = (Q.sum() - want.sum()) / math.sqrt(var.sum())
    return z, ztot, Q, want


def test_optically_thin_closed_form_through_the_estimator(oracle, tables):
    for name, wkey in (("uneven", "wq"), ("even", "wq_asym")):
        z, ztot, Q, want = _thin_z(oracle, tables[name], tables[wkey])
        assert np.max(np.abs(z)) < 4.0, (name, wkey, np.max(np.abs(z)))
        assert abs(ztot) < 4.0, (name, wkey, ztot)
        # the check has power: the per-cell standard error is well below 1 %
        assert np.median(np.abs(Q - want) / np.abs(want)) < 0.01


@pytest.mark.parametrize("mistake,table,wkey", [(1, "uneven", "wq"), (2, "even", "wq_asym"), (3, "uneven", "wq")])
def test_thin_closed_form_fails_on_planted_mistakes(oracle, tables, mistake, table, wkey):
    oracle.plant_mistake(mistake)
    try:
        z, ztot, _, _ = _thin_z(oracle, tables[table], tables[wkey])
    finally:
        oracle.plant_mistake(0)
    assert abs(ztot) > 8.0, (mistake, ztot)
