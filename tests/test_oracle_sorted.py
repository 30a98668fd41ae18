"""CPU pins of the oracle's sorted compat mode (srt == 1: code/mc_gpu.cu:941-957
sort_idx, :786-812 kernel_find, :897-930 wave_find).

The oracle restates the reference's count-and-replay loops.  They are pinned
here by what they must produce, computed another way:
- the band order is a stable argsort of kavg;
- with Sobol directions the slot's XORWOW stream feeds only the (band, g)
  draws, exactly as in the unsorted mode, so the sorted bundles are the
  unsorted draws re-ordered by (band rank, g), with the unsorted mode's
  directions left in place (Sigma counts = photmax, same multiset);
- with XORWOW directions the direction of bundle h is uniforms 2n+2h (phi)
  and 2n+2h+1 (theta) of the stream, after the 2n (band, g) draws.
"""
import numpy as np
import pytest

from dnsmc_amd import workloads


def _gas_problem(oracle, nB=24, nQ=4, R=3000):
    t = workloads.synth_ck_table(5, nB=nB, nQ=nQ)
    Tmax = 1010.0
    ref = oracle.build_proposal(1, t, Tmax, R, np.ones(nQ))
    fx, cx = oracle.grid("uniform", 2, 1.0)
    fy, cy = oracle.grid("tanh", 2, 2.0, beta=1.5)
    fz, cz = oracle.grid("uniform", 2, 1.0)
    T = np.array([900.0, 950.0, 1000.0, 1010.0, 880.0, 990.0, 930.0, 970.0], np.float32)
    Tnb = np.asarray(t["Tnb"], np.float32)
    inv_dT = np.float32(1.0) / (Tnb[1] - Tnb[0])
    d = dict(levels=[dict(n=[2, 2, 2], faces=[fx, fy, fz], T=T, maxs=1000, centres=[cx, cy, cz])],
             bc=[[1, 1], [2, 2], [1, 1]], polar_axis=1, tau_cut=6.9, spectral=0, nB=nB, nQ=nQ, nT=len(Tnb),
             inv_dT=float(inv_dT), mT0=float(-Tnb[0] * inv_dT), kq=ref["kq"], Ib=ref["Ib"],
             Ibw=np.zeros(6 * nB, np.float32), cdf_band=ref["cdf_band"], cdf_g=ref["cdf_g"],
             inv_emax=ref["inv_emax"], Ibmax=ref["Ibmax_d"], Ibmax_d=ref["Ibmax_d"], Tmax=Tmax, g_search=1,
             u_off=1.0)
    kavg = np.asarray(t["kavg"][:, -1], np.float32)
    return oracle.problem(d), kavg


def _band_g(descs, nQ):
    band = (descs["w0"] >> 24) | (((descs["w1"] >> 24) & 3) << 8)
    g = descs["w1"] >> 26
    return band.astype(np.int64), g.astype(np.int64)


def test_sort_idx_is_stable_argsort(oracle):
    rng = np.random.default_rng(3)
    kavg = rng.choice(np.array([0.5, 1.0, 2.0, 3.5], np.float32), 40)      # many ties
    np.testing.assert_array_equal(oracle.sort_idx(kavg), np.argsort(kavg, kind="stable"))
    kavg = rng.random(100).astype(np.float32)
    np.testing.assert_array_equal(oracle.sort_idx(kavg), np.argsort(kavg, kind="stable"))


@pytest.mark.parametrize("cell,subseq,v", [(0, 7, 0), (5, 123456, 2)])
def test_sorted_sobol_is_the_unsorted_draws_in_band_rank_order(oracle, cell, subseq, v):
    R = 3000
    P, kavg = _gas_problem(oracle, R=R)
    order = oracle.sort_idx(kavg)
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    un = oracle.gen_compat(P, 1234, subseq, 1, v * R, cell, R)
    so = oracle.gen_compat_sorted(P, 1234, subseq, 1, v * R, cell, R, order)
    bu, gu = _band_g(un, 4)
    bs, gs = _band_g(so, 4)
    key_u = rank[bu] * 4 + gu
    key_s = rank[bs] * 4 + gs
    np.testing.assert_array_equal(key_s, np.sort(key_u))                    # same multiset, replay order
    assert np.all(np.diff(key_s) >= 0)
    # directions: the Sobol points stay where they were
    np.testing.assert_array_equal(so["w0"] & 0xFFFFFF, un["w0"] & 0xFFFFFF)
    np.testing.assert_array_equal(so["w1"] & 0xFFFFFF, un["w1"] & 0xFFFFFF)
    # weights belong to (band, g) and the cell: equal to an unsorted bundle's of the same bin
    wmap = {}
    for k, w, wi in zip(key_u, un["w"], un["winv"]):
        wmap[int(k)] = (w, wi)
    for k, w, wi in zip(key_s, so["w"], so["winv"]):
        assert wmap[int(k)] == (w, wi)


def test_sorted_xorwow_directions_follow_the_band_draws(oracle):
    R = 1000
    P, kavg = _gas_problem(oracle, R=R)
    order = oracle.sort_idx(kavg)
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    so = oracle.gen_compat_sorted(P, 1234, 99, 0, 0, 3, R, order)
    # the (band, g) draws are the first 2R uniforms of the stream: the same as
    # the unsorted Sobol mode's, whose XORWOW use is exactly those draws
    un = oracle.gen_compat(P, 1234, 99, 1, 0, 3, R)
    bu, gu = _band_g(un, 4)
    bs, gs = _band_g(so, 4)
    np.testing.assert_array_equal(rank[bs] * 4 + gs, np.sort(rank[bu] * 4 + gu))
    st = oracle.xorwow_state(1234, 99, 0)
    u = oracle.xorwow_draws(st, 4 * R)
    np.testing.assert_array_equal(so["w1"] & 0xFFFFFF, u[2 * R::2] >> 8)       # phi first
    np.testing.assert_array_equal(so["w0"] & 0xFFFFFF, u[2 * R + 1::2] >> 8)   # then theta
