"""CPU pins of the oracle's random-number and sampling primitives.

Philox4x32-10 against the Random123 known-answer vectors; XORWOW jumps against
brute-force stepping; Sobol32 dims 0/1 against scipy's Joe-Kuo generator; the
reference's CDF search (code/mc_gpu.cu:842-853, :881-893) against a linear scan.
"""
import math

import numpy as np
import pytest


# Random123 kat_vectors, philox4x32 10 rounds: (ctr, key, expected)
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, want):
    assert oracle.philox(ctr, key).tolist() == want


def _xorwow_py(x, d, n):
    """Marsaglia (2003), J. Stat. Softw. 8(14) p.5, 'xorwow' (d += 362437)."""
    x = list(x)
    out = []
    for _ in range(n):
        t = x[0] ^ (x[0] >> 2)
        x = x[1:] + [x[4] ^ ((x[4] << 4) & 0xFFFFFFFF) ^ t ^ ((t << 1) & 0xFFFFFFFF)]
        d = (d + 362437) & 0xFFFFFFFF
        out.append((d + x[4]) & 0xFFFFFFFF)
    return out


def test_xorwow_recurrence_matches_published(oracle):
    s = oracle.xorwow_state(20240601)
    x0, d0 = list(s.x), s.d
    got = oracle.xorwow_draws(s, 2000).tolist()
    assert got == _xorwow_py(x0, d0, 2000)


@pytest.mark.parametrize("offset", [1, 2, 3, 7, 64, 1000, 65537, 200003])
def test_xorwow_offset_jump_equals_stepping(oracle, offset):
    base = oracle.xorwow_state(1234, 0, 0)
    oracle.xorwow_draws(base, offset)               # brute force: draw `offset` numbers
    jumped = oracle.xorwow_state(1234, 0, offset)
    assert list(jumped.x) == list(base.x) and jumped.d == base.d
    assert oracle.xorwow_draws(jumped, 16).tolist() == oracle.xorwow_draws(base, 16).tolist()


def test_xorwow_subsequence_then_offset(oracle):
    # offset composes with the subsequence jump: (sub, off) == (sub, 0) + off draws
    for sub in (1, 5, 123456789):
        a = oracle.xorwow_state(99, sub, 0)
        oracle.xorwow_draws(a, 333)
        b = oracle.xorwow_state(99, sub, 333)
        assert list(a.x) == list(b.x) and a.d == b.d
    # distinct subsequences give distinct streams
    s = {tuple(oracle.xorwow_draws(oracle.xorwow_state(7, k, 0), 4)) for k in range(64)}
    assert len(s) == 64


def test_sobol_dims_match_joe_kuo(oracle):
    from scipy.stats import qmc
    n = 4096
    pts = qmc.Sobol(d=2, scramble=False).random(n)
    want = np.round(pts * 2.0 ** 32).astype(np.uint64)
    for dim in (0, 1):
        got = oracle.sobol_points(dim, 0, n).astype(np.uint64)
        np.testing.assert_array_equal(got, want[:, dim])


def test_sobol_net_property(oracle):
    # (0, m, 2)-net: every elementary interval of area 2^-m holds one point
    m = 10
    a = oracle.sobol_points(0, 0, 1 << m).astype(np.int64)
    b = oracle.sobol_points(1, 0, 1 << m).astype(np.int64)
    for k in range(m + 1):
        ia = a >> (32 - k) if k else np.zeros_like(a)
        ib = b >> (32 - (m - k)) if m - k else np.zeros_like(b)
        cells = ia * (1 << (m - k)) + ib
        assert len(np.unique(cells)) == 1 << m


def test_uniform_conversions(oracle):
    assert oracle.lib.orc_u01_incl(0) == pytest.approx(2.0 ** -24, rel=0)
    assert oracle.lib.orc_u01_incl(0xFFFFFFFF) == 1.0
    # open conversion: (m + 1/2) 2^-23 on the top 23 bits, exact in FP32
    assert oracle.lib.orc_u01_open(0) == 2.0 ** -24
    assert oracle.lib.orc_u01_open(0xFFFFFFFF) == 1.0 - 2.0 ** -24
    rng = np.random.default_rng(1)
    xs = rng.integers(0, 2 ** 32, 2000, dtype=np.uint64)
    for x in xs:
        u = oracle.lib.orc_u01_open(int(x))
        assert u == ((int(x) >> 9) + 0.5) * 2.0 ** -23


def test_sincos2pi_accuracy(oracle):
    rng = np.random.default_rng(5)
    us = np.concatenate([rng.random(20000, dtype=np.float32), np.float32([0, 0.25, 0.5, 0.75, 1 - 2 ** -24])])
    err = 0.0
    for u in us:
        s, c = oracle.sincos2pi(float(u))
        a = 2.0 * math.pi * float(u)
        err = max(err, abs(s - math.sin(a)), abs(c - math.cos(a)))
    assert err < 4e-7


def _scan(cdf, R):
    if R == 1.0:
        return len(cdf) - 1
    return min(int(np.sum(cdf <= R)), len(cdf) - 1)


def test_cdf_search_equals_linear_scan(oracle):
    rng = np.random.default_rng(3)
    for trial in range(200):
        n = int(rng.integers(1, 130))
        p = rng.random(n).astype(np.float32) ** 3
        p[rng.random(n) < 0.2] = 0.0                      # empty bins -> ties in the CDF
        c = np.cumsum(p, dtype=np.float64)
        c = (c / c[-1] if c[-1] > 0 else np.linspace(0, 1, n)).astype(np.float32)
        Rs = np.concatenate([rng.random(50).astype(np.float32), c, np.float32([0.0, 1.0, 2.0 ** -24])])
        for R in Rs:
            assert oracle.cdf_search(c, float(R)) == _scan(c, np.float32(R)), (trial, R)
