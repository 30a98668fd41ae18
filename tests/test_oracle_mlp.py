"""CPU pins of the proposal network's FP64 reference (oracle/mlp.c):
analytic gradients against central finite differences of its own loss, the
forward pass against a numpy restatement of the layer algebra, and the loss
at a known point (uniform logits: log ndir + log nbins)."""
import math

import numpy as np
import pytest


def _net(rng, nin=6, nh=8, nout=13):
    W1 = rng.standard_normal((nin, nh)) * 0.5
    b1 = rng.standard_normal(nh) * 0.1
    W2 = rng.standard_normal((nh, nh)) * 0.5
    b2 = rng.standard_normal(nh) * 0.1
    W3 = rng.standard_normal((nh, nout)) * 0.5
    b3 = rng.standard_normal(nout) * 0.1
    return [W1, b1, W2, b2, W3, b3]


def test_forward_matches_numpy(oracle):
    rng = np.random.default_rng(1)
    p = _net(rng)
    X = rng.standard_normal((5, 6))
    H1, H2, Z = oracle.mlp_forward(X, *p)
    h1 = np.tanh(X @ p[0] + p[1])
    h2 = np.tanh(h1 @ p[2] + p[3])
    np.testing.assert_allclose(H1, h1, rtol=1e-12)
    np.testing.assert_allclose(H2, h2, rtol=1e-12)
    np.testing.assert_allclose(Z, h2 @ p[4] + p[5], rtol=1e-12, atol=1e-14)


def test_loss_at_uniform_logits(oracle):
    rng = np.random.default_rng(2)
    p = _net(rng)
    p[4][:] = 0.0
    p[5][:] = 0.0                                   # all logits zero
    X = rng.standard_normal((7, 6))
    ndir = 5
    td = rng.integers(0, ndir, 7)
    tb = rng.integers(0, 13 - ndir, 7)
    loss, _ = oracle.mlp_loss_grad(X, *p, ndir, td, tb, np.ones(7))
    assert loss == pytest.approx(math.log(ndir) + math.log(13 - ndir), rel=1e-12)


def test_gradients_match_finite_differences(oracle):
    rng = np.random.default_rng(3)
    p = _net(rng)
    nb, ndir = 9, 5
    X = rng.standard_normal((nb, 6))
    td = rng.integers(0, ndir, nb)
    tb = rng.integers(0, 13 - ndir, nb)
    wt = rng.random(nb) + 0.1
    loss, g = oracle.mlp_loss_grad(X, *p, ndir, td, tb, wt)
    eps = 1e-6
    for k, (P, G) in enumerate(zip(p, g)):
        flat = P.reshape(-1)
        gflat = G.reshape(-1)
        for idx in rng.choice(flat.size, size=min(flat.size, 12), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp, _ = oracle.mlp_loss_grad(X, *p, ndir, td, tb, wt)
            flat[idx] = old - eps
            lm, _ = oracle.mlp_loss_grad(X, *p, ndir, td, tb, wt)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert gflat[idx] == pytest.approx(fd, rel=1e-5, abs=1e-9), (k, idx)


def test_dropped_term_is_caught(oracle):
    """Planting a mistake in the oracle's inputs (a transposed W2) changes the
    gradient well beyond the finite-difference tolerance used above."""
    rng = np.random.default_rng(4)
    p = _net(rng)
    X = rng.standard_normal((6, 6))
    td = rng.integers(0, 5, 6)
    tb = rng.integers(0, 8, 6)
    _, g = oracle.mlp_loss_grad(X, *p, 5, td, tb, np.ones(6))
    q = [a.copy() for a in p]
    q[2] = q[2].T.copy()
    _, gq = oracle.mlp_loss_grad(X, *q, 5, td, tb, np.ones(6))
    assert np.max(np.abs(gq[0] - g[0])) > 1e-3 * np.max(np.abs(g[0]))


# ------------------------------------------- BF16 rounding mode (the kernels')
def bf16_numpy(a):
    """FP64 -> FP32 (nearest even) -> BF16 (nearest even on the FP32 bits),
    vectorised on the integer bits."""
    u = np.asarray(a, np.float64).astype(np.float32).view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = np.where(finite, (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000, u & 0xFFFF0000)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def test_bf16_round_matches_numpy(oracle):
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-6, 6, 2000),
                        # exact ties of the FP32 -> BF16 step: round to even both ways
                        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 0.0, -0.0,
                                  2.0 ** -130, 3.4e38, np.inf, -np.inf])])
    got = np.array([oracle.bf16_round(v) for v in x])
    want = bf16_numpy(x)
    np.testing.assert_array_equal(got, want)
    assert oracle.bf16_round(1.0 + 2.0 ** -8) == 1.0                   # tie to even (down)
    assert oracle.bf16_round(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6   # tie to even (up)


def _bf16_numpy_net(X, p, ndir, td, tb, wt):
    """Forward and backward with the kernels' rounding points, in numpy."""
    W1, b1, W2, b2, W3, b3 = p
    H1 = np.tanh(X @ W1 + b1)
    H1q = bf16_numpy(H1)
    H2 = np.tanh(H1q @ W2 + b2)
    H2q = bf16_numpy(H2)
    Z = H2q @ W3 + b3
    sc = wt / wt.sum()
    dZ = np.zeros_like(Z)
    for lo, hi, tgt in ((0, ndir, td), (ndir, Z.shape[1], tb)):
        z = Z[:, lo:hi]
        e = np.exp(z - z.max(1, keepdims=True))
        sm = e / e.sum(1, keepdims=True)
        sm[np.arange(len(tgt)), tgt] -= 1.0
        dZ[:, lo:hi] = sc[:, None] * sm
    dZq = bf16_numpy(dZ)
    dA2 = (dZq @ W3.T) * (1.0 - H2 ** 2)
    dA2q = bf16_numpy(dA2)
    dA1 = (dA2q @ W2.T) * (1.0 - H1 ** 2)
    g = [X.T @ dA1, dA1.sum(0), H1q.T @ dA2q, dA2.sum(0), H2q.T @ dZq, dZ.sum(0)]
    return Z, g


def test_bf16_mode_matches_numpy(oracle):
    rng = np.random.default_rng(8)
    p = _net(rng, nh=16, nout=29)
    nb, ndir = 11, 7
    X = rng.standard_normal((nb, 6))
    td = rng.integers(0, ndir, nb)
    tb = rng.integers(0, 29 - ndir, nb)
    wt = rng.random(nb) + 0.1
    Zn, gn = _bf16_numpy_net(X, p, ndir, td, tb, wt)
    _, _, Z = oracle.mlp_forward(X, *p, bf16=True)
    np.testing.assert_allclose(Z, Zn, rtol=1e-12, atol=1e-13)
    _, g = oracle.mlp_loss_grad(X, *p, ndir, td, tb, wt, bf16=True)
    for a, b in zip(g, gn):
        np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-13)
    # the rounding is there (the modes differ) and stays at BF16 scale
    _, gf = oracle.mlp_loss_grad(X, *p, ndir, td, tb, wt)
    rel = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(g, gf)]
    assert max(rel) > 1e-5 and max(rel) < 2e-2, rel


# ------------------------------------------------ learned draw (orc_learned_draw)
def _rows(rng, nslots, ndir, nbin):
    """Unnormalised inclusive cumulative weights, as the proposal writes them."""
    noutp = ((ndir + nbin + 63) // 64) * 64
    U = np.zeros((nslots, noutp), np.float32)
    for s in range(nslots):
        U[s, :ndir] = np.cumsum(rng.uniform(0.05, 3.0, ndir)).astype(np.float32)
        U[s, ndir:ndir + nbin] = np.cumsum(rng.uniform(0.0, 2.0, nbin) ** 3).astype(np.float32)
    return U


def _plain(rng, nbin):
    p = rng.uniform(0.1, 1.0, nbin)
    p /= p.sum()
    pcdf = np.minimum(np.cumsum(p), 1.0).astype(np.float32)
    pcdf[-1] = 1.0
    return p.astype(np.float32), pcdf


def _decode(w0, w1):
    band = (w0 >> 24) | (((w1 >> 24) & 3) << 8)
    g = w1 >> 26
    return band.astype(np.int64), g.astype(np.int64), (w0 & 0xFFFFFF) >> 21, (w1 & 0xFFFFFF) >> 21


def test_learned_draw_pure_plain_mixture_is_the_plain_search(oracle):
    """eps = 1: the mixture is the plain proposal, so the (band, g) bin is the
    plain CDF search of the same uniform (orc_cdf_search, pinned by a linear
    scan), the direction bin the uniform 64-bin search, and p/q = 1 up to the
    FP32 steps of the CDF."""
    rng = np.random.default_rng(3)
    nQ, nB, R = 4, 5, 3000
    nbin = nB * nQ
    U = _rows(rng, 2, 64, nbin)
    p, pcdf = _plain(rng, nbin)
    cells = np.array([7, 123], np.int32)
    w0, w1, ratio = oracle.learned_draw(U, pcdf, p, cells, R, 64, nQ, 2, 99, 1.0)
    band, g, th, ph = _decode(w0, w1)
    bg = band * nQ + g
    pd = ((np.arange(64) + 1) / 64.0).astype(np.float32)
    for sl, cell in enumerate(cells):
        for r in range(0, R, 97):
            x = oracle.philox([r, cell, 2, 1], [99, 0])
            q = sl * R + r
            assert bg[q] == oracle.cdf_search(pcdf, float(np.float32(oracle.lib.orc_u01_open(int(x[1])))))
            d = oracle.cdf_search(pd, float(np.float32(oracle.lib.orc_u01_open(int(x[0])))))
            assert th[q] == d >> 3 and ph[q] == d & 7
    steps = np.diff(np.concatenate([[0.0], pcdf.astype(np.float64)]))
    np.testing.assert_allclose(ratio, p[bg] / steps[bg], rtol=1e-5)
    np.testing.assert_allclose(ratio, 1.0, rtol=1e-5)


def test_learned_draw_frequencies_follow_the_mixture(oracle):
    """Bin frequencies of many draws against the mixture (1 - eps) q + eps p
    computed independently from the rows (numpy), z < 5 per bin; and
    E_mixture[p/q] = 1."""
    rng = np.random.default_rng(5)
    nQ, nB, R, eps = 3, 4, 400000, 0.3
    nbin = nB * nQ
    U = _rows(rng, 1, 64, nbin)
    p, pcdf = _plain(rng, nbin)
    w0, w1, ratio = oracle.learned_draw(U, pcdf, p, np.array([11], np.int32), R, 64, nQ, 0, 1234, eps)
    band, g, th, ph = _decode(w0, w1)
    for idx, n, Uh, P in ((band * nQ + g, nbin, U[0, 64:64 + nbin], p.astype(np.float64)),
                          (th * 8 + ph, 64, U[0, :64], np.full(64, 1 / 64.0))):
        q = np.diff(np.concatenate([[0.0], Uh.astype(np.float64)])) / float(Uh[-1])
        mix = (1 - eps) * q + eps * P
        cnt = np.bincount(idx, minlength=n)
        z = (cnt - R * mix) / np.sqrt(R * mix * (1 - mix))
        assert np.max(np.abs(z)) < 5.0, z
    assert abs(float(np.mean(ratio.astype(np.float64))) - 1.0) < 5 * float(np.std(ratio)) / math.sqrt(R)
