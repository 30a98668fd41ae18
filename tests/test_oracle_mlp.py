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
