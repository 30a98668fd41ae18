"""GPU tests of the learned proposal: the BF16 MFMA forward and backward
kernels against the FP64 oracle (oracle/mlp.c, itself pinned by finite
differences), training on plain-run tallies, unbiasedness of the p/q
reweighted estimator and its variance against the plain estimator.

Tolerances: BF16 weights are exact inputs to both sides.  The kernels round
the product operands H1, H2 (forward) and dZ, dA2 (backward) to BF16; the
oracle's BF16 mode (orc_mlp_*_bf16) rounds at the same points, so logits and
every gradient block agree with it to 1e-3 in relative norm (what is left is
FP32 accumulation and the rare element whose FP32 and FP64 values round to
different BF16 neighbours), and with the unrounded FP64 network to 1e-2.
"""
import numpy as np
import pytest

import dnsmc_amd
from dnsmc_amd import learned, workloads
from dnsmc_amd.devarray import DeviceArray

pytestmark = pytest.mark.gpu


def t(a):
    return DeviceArray.from_numpy(np.ascontiguousarray(a))


def _random_net(nbin, seed):
    net = learned.ProposalNet(nbin, seed=seed)
    rng = np.random.default_rng(seed + 1)
    th = net.theta.numpy()
    L = net.L
    th[L["W3"]:L["b3"]] = rng.uniform(-1, 1, L["b3"] - L["W3"]) * 0.3     # non-trivial logits
    th[L["b1"]:L["W2"]] = rng.normal(0, 0.1, 64)
    th[L["b3"]:L["P"]] = rng.normal(0, 0.2, L["P"] - L["b3"])
    return learned.ProposalNet(nbin, theta=th)


def _oracle_params(net):
    """The network's parameters as the kernels see them (BF16 weights) in FP64."""
    th = net.theta.numpy().astype(np.float64)
    L, no = net.L, net.L["noutp"]
    W1 = learned.bf16_to_f32(net.W1b.numpy()).reshape(6, 64).astype(np.float64)
    W2 = learned.bf16_to_f32(net.W2b.numpy()).reshape(64, 64).astype(np.float64)
    W3 = learned.bf16_to_f32(net.W3b.numpy()).reshape(64, no).astype(np.float64)
    nout = 64 + net.nbin
    return [W1, th[L["b1"]:L["W2"]], W2, th[L["b2"]:L["W3"]], W3[:, :nout], th[L["b3"]:L["b3"] + nout]]


@pytest.mark.parametrize("nbin", [119, 1600])
def test_forward_matches_oracle(oracle, nbin):
    net = _random_net(nbin, 3)
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (300, 6)).astype(np.float32)
    Z = net.forward(t(X)).numpy()
    nout = 64 + nbin
    p = _oracle_params(net)
    _, _, Zq = oracle.mlp_forward(X.astype(np.float64), *p, bf16=True)
    _, _, Zr = oracle.mlp_forward(X.astype(np.float64), *p)
    relq = np.linalg.norm(Z[:, :nout] - Zq) / np.linalg.norm(Zq)
    relr = np.linalg.norm(Z[:, :nout] - Zr) / np.linalg.norm(Zr)
    assert relq < 1e-3, relq
    assert relr < 1e-2, relr


@pytest.mark.parametrize("nbin", [119, 1600])
def test_gradients_match_oracle(oracle, nbin):
    net = _random_net(nbin, 7)
    rng = np.random.default_rng(9)
    n = 1000
    X = rng.normal(0, 1, (n, 6)).astype(np.float32)
    td = rng.integers(0, 64, n).astype(np.int32)
    tb = rng.integers(0, nbin, n).astype(np.int32)
    wt = rng.random(n).astype(np.float32)
    wt[rng.random(n) < 0.3] = 0.0
    loss = net.loss_grad(t(X), t(td), t(tb), t(wt))
    g = net.grad.numpy()
    p = _oracle_params(net)
    lr, gr = oracle.mlp_loss_grad(X.astype(np.float64), *p, 64, td, tb, wt.astype(np.float64))
    lq, gq = oracle.mlp_loss_grad(X.astype(np.float64), *p, 64, td, tb, wt.astype(np.float64), bf16=True)
    assert loss == pytest.approx(lq, rel=1e-4)
    assert loss == pytest.approx(lr, rel=1e-3)
    L, no = net.L, net.L["noutp"]
    nout = 64 + nbin
    mine = [("W1", g[L["W1"]:L["b1"]]), ("b1", g[L["b1"]:L["W2"]]), ("W2", g[L["W2"]:L["b2"]]),
            ("b2", g[L["b2"]:L["W3"]]), ("W3", g[L["W3"]:L["b3"]].reshape(64, no)[:, :nout].ravel()),
            ("b3", g[L["b3"]:L["b3"] + nout])]
    for (name, got), wq, wr in zip(mine, gq, gr):
        relq = np.linalg.norm(got - wq.ravel()) / np.linalg.norm(wq)
        relr = np.linalg.norm(got - wr.ravel()) / np.linalg.norm(wr)
        assert relq < 1e-3, (name, relq)
        assert relr < 1e-2, (name, relr)
    # padding columns get no gradient
    assert np.all(g[L["W3"]:L["b3"]].reshape(64, no)[:, nout:] == 0.0)


def test_training_fits_a_known_distribution():
    """Samples drawn from a feature-dependent categorical distribution: the
    trained network's softmax approaches it (loss near the entropy)."""
    nbin = 40
    rng = np.random.default_rng(1)
    n = 64 * 1024
    X = rng.normal(0, 1, (n, 6)).astype(np.float32)
    logit_d = np.outer(X[:, 0], np.linspace(-2, 2, 64))
    logit_b = np.outer(X[:, 1], np.linspace(-1.5, 1.5, nbin))
    def draw(lg):
        p = np.exp(lg - lg.max(1, keepdims=True))
        p /= p.sum(1, keepdims=True)
        c = p.cumsum(1)
        u = rng.random((lg.shape[0], 1))
        return (u > c).sum(1).clip(0, lg.shape[1] - 1), p
    td, pd = draw(logit_d)
    tb, pb = draw(logit_b)
    ent = np.mean(-np.sum(pd * np.log(pd + 1e-30), 1) - np.sum(pb * np.log(pb + 1e-30), 1))
    samples = (t(X), t(td.astype(np.int32)), t(tb.astype(np.int32)), t(np.ones(n, np.float32)))
    net = learned.ProposalNet(nbin, seed=2)
    hist = learned.train(net, samples, epochs=6, batch=4096, lr=3e-3, seed=3)
    assert hist[-1] < hist[0]
    assert hist[-1] < ent + 0.05 * ent, (hist, ent)


def _small_case(table_dir, mode, R, tally=0, n=(16, 24, 16)):
    return workloads.channel_config(n, rays_per_cell=R, spectral=0, Tw=573.0, table_dir=table_dir,
                                    mode=mode, tally=tally)


def _field(seed, n):
    return workloads.ghosted(workloads.field_fourier(seed, n))


@pytest.fixture(scope="module")
def trained(ck_table_dir):
    """A network trained on plain tallies of three snapshots of a small channel."""
    tdir, _ = ck_table_dir
    n = (16, 24, 16)
    samples = learned.SampleSet()
    for seed in (31, 32, 33):
        S = dnsmc_amd.Solver(_small_case(tdir, 0, 32, n=n))
        S.load_tables()
        S.set_record(True)
        S.set_field(0, _field(seed, n))
        S.run()
        S.wait()
        samples.add_run(S)
        nbin = int(S.array("scalars")[3]) * int(S.array("scalars")[4])
        S.close()
    net = learned.ProposalNet(nbin, seed=0)
    hist = learned.train(net, samples.cat(), epochs=4, batch=8192, lr=2e-3, seed=0)
    return net, hist, tdir, n


def _run(tdir, n, mode, R, field, net=None, seed=20240601):
    cfg = _small_case(tdir, mode, R, n=n)
    cfg.seed = seed
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    if net is not None:
        S.set_mlp(*net.host_weights(), mix_eps=0.1)
    S.set_field(0, field)
    S.run()
    out = S.results(0)
    S.close()
    return out


def test_training_on_tallies_reduces_loss(trained):
    _, hist, _, _ = trained
    assert hist[-1] < hist[0]


def test_learned_estimator_is_unbiased_and_reduces_variance(trained):
    """Held-out snapshot, equal ray budget.  Per-cell z-scores are not used
    as the bar: the plain estimator is heavy-tailed (rare long rays to the cold
    walls), so its own per-cell variance estimate is occasionally far too small
    at 256 rays.  Sums over cells are Gaussian: the bar is the z of the total
    and of every wall-normal row (256 cells each), and the spread of the
    per-cell z."""
    net, _, tdir, n = trained
    field = _field(51, n)                                    # held-out snapshot
    plain = _run(tdir, n, 0, 256, field, seed=101)
    lrn = _run(tdir, n, 2, 256, field, net=net, seed=202)
    var_p, var_l = plain["sdev"] ** 2, lrn["sdev"] ** 2
    dq = lrn["Q"] - plain["Q"]
    z = dq / np.sqrt(var_p + var_l)
    assert abs(np.mean(z)) < 0.1 and np.std(z) < 1.3, (np.mean(z), np.std(z))
    ztot = np.sum(dq) / np.sqrt(np.sum(var_p + var_l))
    assert abs(ztot) < 4.0, ztot
    rows = dq.reshape(n[2], n[1], n[0]).sum(axis=(0, 2))
    rvar = (var_p + var_l).reshape(n[2], n[1], n[0]).sum(axis=(0, 2))
    zrow = rows / np.sqrt(rvar)
    assert np.max(np.abs(zrow)) < 4.5, zrow
    ratio = np.sum(var_p) / np.sum(var_l)
    assert ratio > 3.0, ratio


def test_learned_estimator_unbiased_with_coarsening(trained):
    """The learned generator with two coarsening levels (the batched traversal
    through the coarse field): the p/q-reweighted estimator agrees with the
    plain one on the same configuration (any proposal is unbiased)."""
    net, _, tdir, n = trained
    field = _field(52, n)
    kw = dict(grid_num=2, maxs=(4, 100000, 100000, 100000, 100000, 100000))
    out = []
    for mode, seed, W in ((0, 303, None), (2, 404, net)):
        cfg = workloads.channel_config(n, rays_per_cell=256, spectral=0, Tw=573.0, table_dir=tdir, mode=mode, **kw)
        cfg.seed = seed
        S = dnsmc_amd.Solver(cfg)
        S.load_tables()
        if W is not None:
            S.set_mlp(*W.host_weights(), mix_eps=0.1)
        S.set_field(0, field)
        S.run()
        out.append(S.results(0))
        S.close()
    plain, lrn = out
    var = plain["sdev"] ** 2 + lrn["sdev"] ** 2
    dq = lrn["Q"] - plain["Q"]
    ztot = np.sum(dq) / np.sqrt(np.sum(var))
    assert abs(ztot) < 4.0, ztot
    rows = dq.reshape(n[2], n[1], n[0]).sum(axis=(0, 2)) / np.sqrt(var.reshape(n[2], n[1], n[0]).sum(axis=(0, 2)))
    assert np.max(np.abs(rows)) < 4.5, rows
    assert np.sum(plain["sdev"] ** 2) / np.sum(lrn["sdev"] ** 2) > 2.0


@pytest.mark.parametrize("nsnap,pipes", [(3, None), (1, None), (3, "1")])
def test_snapshotwise_learned_matches_batched(trained, monkeypatch, nsnap, pipes):
    """Learned solve of one or three snapshots: descriptors of a whole
    snapshot and one traversal launch per snapshot (generation of the next
    overlapped) against the batched pipelines: the same rays, so equal step
    totals and tallies equal up to the order of the FP64 atomic additions.
    DNSMC_PIPES=1 restricts the batched loop only; the snapshot-wise path
    keeps its two streams."""
    net, _, tdir, n = trained
    if pipes is not None:
        monkeypatch.setenv("DNSMC_PIPES", pipes)
    monkeypatch.setenv("DNSMC_LEARNED_FUSED", "0")
    out = {}
    for sw in ("1", "0"):
        monkeypatch.setenv("DNSMC_SNAPWISE", sw)
        cfg = workloads.channel_config(n, rays_per_cell=64, spectral=0, Tw=573.0, table_dir=tdir, mode=2, nsnap=nsnap,
                                       batch_rays=64 * 1000)
        S = dnsmc_amd.Solver(cfg)
        S.load_tables()
        S.set_mlp(*net.host_weights(), mix_eps=0.1)
        for s in range(nsnap):
            S.set_field(s, _field(70 + s, n))
        S.run()
        st = S.stats()
        out[sw] = (st, [S.results(s) for s in range(nsnap)])
        S.close()
    (s1, r1), (s0, r0) = out["1"], out["0"]
    assert s1["n_trace_launches"] == nsnap and s0["n_trace_launches"] > nsnap
    assert s1["n_steps"] == s0["n_steps"]
    for s in range(nsnap):
        np.testing.assert_allclose(r1[s]["E"], r0[s]["E"], rtol=1e-12, atol=0)
        np.testing.assert_allclose(r1[s]["A"], r0[s]["A"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("R,slots", [(40, None), (40, "0"), (40, "3"), (256, None)])
def test_learned_draw_matches_oracle(oracle, trained, monkeypatch, R, slots):
    """The learned generator ray by ray against the oracle's learned draw
    (orc_learned_draw, pinned on the CPU by the plain limit and the mixture
    frequencies) on the same cumulative-weight rows: descriptor words bit for
    bit and p/q to the last bit (the same FP32 operations).  By default 40
    rays per cell stage four cells' rows per block in LDS and 256 three
    (raygen_learned_lds); DNSMC_RAYGEN_SLOTS=0 forces the search of the rows
    in global memory (raygen_learned), =3 three cells per block at 40."""
    net, _, tdir, n = trained
    monkeypatch.setenv("DNSMC_SNAPWISE", "0")
    monkeypatch.setenv("DNSMC_LEARNED_FUSED", "0")
    if slots:
        monkeypatch.setenv("DNSMC_RAYGEN_SLOTS", slots)
    cfg = _small_case(tdir, 2, R, n=n)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    S.set_mlp(*net.host_weights(), mix_eps=0.1)
    S.set_field(0, _field(77, n))
    S.run()
    S.wait()
    ncell = n[0] * n[1] * n[2]
    rays = S.rays(ncell * R)
    U = S.array("qcdf").reshape(ncell, -1)
    p_bg = S.array("p_bg")
    pcdf = S.array("pcdf")
    nQ = int(S.array("scalars")[4])
    S.close()
    w0, w1, ratio = oracle.learned_draw(U, pcdf, p_bg, np.arange(ncell, dtype=np.int32), R, 64, nQ, 0,
                                        cfg.seed, np.float32(0.1))
    np.testing.assert_array_equal(rays["w0"], w0)
    np.testing.assert_array_equal(rays["w1"], w1)
    np.testing.assert_array_equal(rays["w"], ratio)


def test_proposal_rows_match_the_network(trained, monkeypatch):
    """mlp_proposal's cumulative weights per cell against the network forward
    (mlp_forward_rows, pinned by test_forward_matches_oracle) on the same
    features: each head's row is the running sum of exp(clamp(z, -60, 60)),
    non-decreasing."""
    net, _, tdir, n = trained
    monkeypatch.setenv("DNSMC_SNAPWISE", "0")
    monkeypatch.setenv("DNSMC_LEARNED_FUSED", "0")
    cfg = _small_case(tdir, 2, 8, n=n)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    S.set_mlp(*net.host_weights(), mix_eps=0.1)
    S.set_field(0, _field(78, n))
    S.run()
    S.wait()
    ncell = n[0] * n[1] * n[2]
    U = S.array("qcdf").reshape(ncell, -1).astype(np.float64)
    X = S.array("features").reshape(ncell, 6)
    S.close()
    Z = net.forward(t(X)).numpy().astype(np.float64)
    nbin = net.nbin
    e = np.exp(np.clip(Z[:, :64 + nbin], -60.0, 60.0))
    for lo, hi in ((0, 64), (64, 64 + nbin)):
        want = np.cumsum(e[:, lo:hi], axis=1)
        np.testing.assert_allclose(U[:, lo:hi], want, rtol=2e-4, atol=1e-30)
        # non-decreasing, so that each bin's draw probability is its step
        assert np.all(np.diff(U[:, lo:hi].astype(np.float32), axis=1) >= 0)


@pytest.mark.parametrize("R", [40, 130])
def test_fused_learned_draw_matches_oracle(oracle, trained, R, monkeypatch):
    """The learned draw inside the traversal (DNSMC_LEARNED_FUSED=1): the
    proposal rows the waves evaluated and every drawn ray are logged, and the
    oracle's learned draw (orc_learned_draw) on those rows gives the same
    descriptor words and p/q bit for bit.  R = 40 puts up to three cells in
    one 64-ray queue draw, R = 130 splits cells over draws and waves."""
    monkeypatch.setenv("DNSMC_LEARNED_FUSED", "1")
    net, _, tdir, n = trained
    cfg = _small_case(tdir, 2, R, n=n)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    S.set_mlp(*net.host_weights(), mix_eps=0.1)
    S.set_field(0, _field(77, n))
    S.set_draw_log(True)
    S.run()
    st = S.stats()
    assert st["n_trace_launches"] == 1 and st["ms_raygen"] == 0.0   # no generator kernel
    ncell = n[0] * n[1] * n[2]
    rays = S.rays(ncell * R)
    U = S.array("lrows").reshape(ncell, -1)
    p_bg = S.array("p_bg")
    pcdf = S.array("pcdf")
    nQ = int(S.array("scalars")[4])
    S.close()
    w0, w1, ratio = oracle.learned_draw(U, pcdf, p_bg, np.arange(ncell, dtype=np.int32), R, 64, nQ, 0,
                                        cfg.seed, np.float32(0.1))
    np.testing.assert_array_equal(rays["w0"], w0)
    np.testing.assert_array_equal(rays["w1"], w1)
    np.testing.assert_array_equal(rays["w"], ratio)


def test_fused_learned_rows_match_the_network(oracle, trained, monkeypatch):
    """The rows the traversal's waves evaluate (BF16 pair products in FP32,
    H1 and H2 rounded to BF16) against the oracle's network in its BF16
    rounding mode on the same features: the running sums of
    exp(clamp(z, -60, 60)) of each head."""
    monkeypatch.setenv("DNSMC_LEARNED_FUSED", "1")
    net, _, tdir, n = trained
    cfg = _small_case(tdir, 2, 8, n=n)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    S.set_mlp(*net.host_weights(), mix_eps=0.1)
    S.set_field(0, _field(78, n))
    S.set_draw_log(True)
    S.run()
    S.wait()
    ncell = n[0] * n[1] * n[2]
    U = S.array("lrows").reshape(ncell, -1).astype(np.float64)
    X = S.array("features").reshape(ncell, 6).astype(np.float64)
    S.close()
    _, _, Z = oracle.mlp_forward(X, *_oracle_params(net), bf16=True)
    nbin = net.nbin
    e = np.exp(np.clip(Z, -60.0, 60.0))
    for lo, hi in ((0, 64), (64, 64 + nbin)):
        want = np.cumsum(e[:, lo:hi], axis=1)
        rel = np.abs(U[:, lo:hi] - want) / want
        assert np.median(rel) < 1e-5 and np.max(rel) < 1e-2, (np.median(rel), np.max(rel))


def test_fused_learned_split_launches_and_snapshots(trained, monkeypatch):
    """Three snapshots in one fused learned launch against the same solve
    split into launches of 37 slots: every ray is drawn from its (ray, cell,
    snapshot) alone, so the step totals are equal and the tallies agree up to
    the order of the FP64 atomic additions."""
    monkeypatch.setenv("DNSMC_LEARNED_FUSED", "1")
    net, _, tdir, n = trained
    out = []
    for slots in (None, "37"):
        if slots:
            monkeypatch.setenv("DNSMC_FUSED_SLOTS", slots)
        cfg = workloads.channel_config(n, rays_per_cell=48, spectral=0, Tw=573.0, table_dir=tdir, mode=2, nsnap=3)
        S = dnsmc_amd.Solver(cfg)
        S.load_tables()
        S.set_mlp(*net.host_weights(), mix_eps=0.1)
        for s in range(3):
            S.set_field(s, _field(90 + s, n))
        S.run()
        st = S.stats()
        out.append((st, [S.results(s) for s in range(3)]))
        S.close()
    (s1, r1), (s2, r2) = out
    assert s1["n_trace_launches"] == 1 and s2["n_trace_launches"] > 1
    assert s1["n_steps"] == s2["n_steps"]
    for s in range(3):
        np.testing.assert_allclose(r1[s]["E"], r2[s]["E"], rtol=1e-12, atol=0)
        np.testing.assert_allclose(r1[s]["A"], r2[s]["A"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("coarse", [False, True])
def test_fused_learned_agrees_with_separate_generation(trained, monkeypatch, coarse):
    """The draw inside the traversal against the default generation kernels
    on the same seed, on the fine grid and with two coarsening levels: both
    draw ray (r, cell, snapshot) from Philox (r, cell, snapshot, 1) on the
    same network, so the rays differ only where the two row evaluations
    (per-wave BF16 pair products against 16x16x16 MFMA tiles) round a bin
    edge differently; the step totals agree to 1e-3 and the per-cell Q to
    within one standard deviation (a few cells hold a ray drawn into a
    neighbouring bin, the rest are identical)."""
    net, _, tdir, n = trained
    kw = dict(grid_num=2, maxs=(4, 100000, 100000, 100000, 100000, 100000)) if coarse else {}
    out = []
    for env in ("0", "1"):
        monkeypatch.setenv("DNSMC_LEARNED_FUSED", env)
        S = dnsmc_amd.Solver(workloads.channel_config(n, rays_per_cell=64, spectral=0, Tw=573.0, table_dir=tdir,
                                                      mode=2, **kw))
        S.load_tables()
        S.set_mlp(*net.host_weights(), mix_eps=0.1)
        S.set_field(0, _field(95, n))
        S.run()
        out.append((S.stats(), S.results(0)))
        S.close()
    (s0, r0), (s1, r1) = out
    assert s1["n_trace_launches"] == 1 and s1["ms_raygen"] == 0.0 and s0["ms_raygen"] > 0.0
    assert s0["n_rays"] == s1["n_rays"]
    assert abs(s0["n_steps"] / s1["n_steps"] - 1.0) < 1e-3
    z = (r0["Q"] - r1["Q"]) / np.sqrt(r0["sdev"] ** 2 + r1["sdev"] ** 2 + 1e-300)
    assert np.max(np.abs(z)) < 1.0, np.max(np.abs(z))
    assert float(np.std(z)) < 0.05, float(np.std(z))


def test_learned_estimator_unbiased_on_the_full_channel(trained):
    """The learned generator at the throughput workload's full shape
    (384x256x256, two snapshots, 8 rays per cell), with the network of the
    small channel (its features do not depend on the grid): the p/q-reweighted
    sums over every wall-normal row of both snapshots agree with the plain
    estimator's (|z| < 4.5 over 512 rows, and in total), and the variance
    drops."""
    net, _, tdir, _ = trained
    n = (384, 256, 256)
    nsnap = 2
    fields = workloads.workload_fields("ck384x8", nsnap)
    out = []
    for mode, seed, W in ((0, 303, None), (2, 404, net)):
        cfg = workloads.channel_config(n, rays_per_cell=8, spectral=0, Tw=573.0, table_dir=tdir, mode=mode,
                                       nsnap=nsnap)
        cfg.seed = seed
        S = dnsmc_amd.Solver(cfg)
        S.load_tables()
        if W is not None:
            S.set_mlp(*W.host_weights(), mix_eps=0.05)
        for s in range(nsnap):
            S.set_field(s, workloads.ghosted(fields[s]))
        S.run()
        out.append([S.results(s) for s in range(nsnap)])
        S.close()
    zmax, num, den, vp, vl = 0.0, 0.0, 0.0, 0.0, 0.0
    for s in range(nsnap):
        p, l = out[0][s], out[1][s]
        var = (p["sdev"] ** 2 + l["sdev"] ** 2).reshape(n[2], n[1], n[0])
        dq = (l["Q"] - p["Q"]).reshape(n[2], n[1], n[0])
        rows = dq.sum(axis=(0, 2)) / np.sqrt(var.sum(axis=(0, 2)))
        zmax = max(zmax, float(np.max(np.abs(rows))))
        num += float(dq.sum())
        den += float(var.sum())
        vp += float(np.sum(p["sdev"] ** 2))
        vl += float(np.sum(l["sdev"] ** 2))
    print("full channel learned vs plain: max row |z| %.2f, total z %.2f, variance ratio %.2f"
          % (zmax, num / np.sqrt(den), vp / vl), flush=True)
    assert zmax < 4.5
    assert abs(num / np.sqrt(den)) < 4.0
    assert vp / vl > 2.0
