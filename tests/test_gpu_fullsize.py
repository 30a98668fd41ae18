"""The product path at the throughput workload's full shape: the 384x256x256
channel, 8 snapshots (profile + Fourier modes) in one fused launch with the
rays drawn in the traversal kernel and the per-wave tally ring, 4 rays per
cell.  Checked on sampled emitting cells of every snapshot against the
oracle (the same Philox draws, code/mc_gpu.cu:813-896, traced cell by cell,
code/mc_gpu.cu:228-438), and through what holds at any size: every ray is
counted, every cell's E is positive and finite, and gas and walls at one
temperature give Q == 0 to rounding (reciprocity) on the full grid.

Bars as in the small-grid parity tests: FP64 cell tallies E, A within 1e-5
relative and Q = E - A within 1e-5 * E; the isothermal |Q| / E below 1e-4 in
every cell (FP32 traversal rounding over up to ~1000 segments; 1.3e-5 is
measured) and below 2e-6 in 99.99 % of them (6.7e-7 measured)."""
import numpy as np
import pytest

import dnsmc_amd
from dnsmc_amd import study, workloads
from helpers import oracle_problem_dict

pytestmark = pytest.mark.gpu

R = 4
NSNAP = 8


def _sample(n, per_snap, seed):
    """Emitting cells of one snapshot: both wall layers, the centreline and
    random interior cells (x fastest, cell = i + n0 * (j + n1 * k))."""
    rng = np.random.default_rng(seed)
    js = [0, n[1] - 1, n[1] // 2] + list(rng.integers(0, n[1], per_snap - 3))
    out = []
    for j in js:
        i, k = int(rng.integers(0, n[0])), int(rng.integers(0, n[2]))
        out.append(i + n[0] * (int(j) + n[1] * k))
    return out


def test_full_channel_eight_snapshots(oracle):
    fields = workloads.workload_fields("ck384x8", NSNAP)
    S, cfg = study.make_solver("ck384x8", R, nsnap=NSNAP, fields=fields)
    n = tuple(cfg.n)
    nc = n[0] * n[1] * n[2]
    S.run()
    st = S.stats()
    print("full channel: %d rays, %d segments, %d trace launch(es), %.1f ms"
          % (st["n_rays"], st["n_steps"], st["n_trace_launches"], st["ms_trace"]), flush=True)
    assert st["n_rays"] == nc * R * NSNAP
    assert st["n_trace_launches"] == 1          # all snapshots in one launch
    d = oracle_problem_dict(S, cfg, 0)
    Tall = S.array("T", 0)
    assert Tall.size == nc * NSNAP
    for s in range(NSNAP):
        res = S.results(s)
        assert np.all(np.isfinite(res["Q"])) and np.all(res["E"] > 0)
        d["levels"][0]["T"] = Tall[s * nc:(s + 1) * nc]
        P = oracle.problem(d)
        for cell in _sample(n, 6, 100 + s):
            descs = oracle.gen_philox(P, cfg.seed, s, cell, 0, R)
            E, A, _, _ = oracle.trace_descs(P, cell, descs)
            assert abs(res["E"][cell] - E) <= 1e-5 * abs(E), (s, cell)
            assert abs(res["A"][cell] - A) <= 1e-5 * abs(A), (s, cell)
            assert abs(res["Q"][cell] - (E - A)) <= 1e-5 * abs(E), (s, cell)
    S.close()
    # gas and walls at one temperature, a node of the table's 25 K grid (the
    # gas's Ib is interpolated in T, the walls' evaluated exactly)
    iso = workloads.ghosted(np.full(n, 575.0, np.float32))
    cfg2 = workloads.channel_config(n, rays_per_cell=R, spectral=0, Tw=575.0, nsnap=2)
    S2 = dnsmc_amd.Solver(cfg2)
    S2.set_tables(workloads.table_arrays(workloads.synth_ck_table(11)))
    for s in range(2):
        S2.set_field(s, iso)
    S2.run()
    print("tau_cut %.6g" % S2.array("scalars")[10], flush=True)
    for s in range(2):
        r = S2.results(s)
        rel = np.abs(r["Q"]) / r["E"]
        m = int(np.argmax(rel))
        print("isothermal snapshot %d: max |Q|/E %.3g at (i, j, k) = (%d, %d, %d), 99.99%% %.3g, above 1e-5: %d"
              % (s, rel[m], m % n[0], (m // n[0]) % n[1], m // (n[0] * n[1]), np.quantile(rel, 0.9999),
                 np.count_nonzero(rel > 1e-5)), flush=True)
        # each ray's transmissivity is a product over its segments in FP32:
        # rounding grows like (segments) * 2^-24, up to 6e-5 for the ~1000
        # segments of the longest rays here, 1e-7 for typical cells
        assert rel[m] <= 1e-4 and np.quantile(rel, 0.9999) <= 2e-6, s
    S2.close()


def test_full_channel_coarsened(oracle):
    """The reference's default coarsening (5 levels, maxs 5/5/5/5/1e5,
    code/param.h:18-35) at the full shape, two snapshots in the fused launch:
    sampled cells against the oracle tracing through the same levels."""
    nsnap = 2
    fields = workloads.workload_fields("ck384x8", nsnap)
    S, cfg = study.make_solver("ck384x8", R, nsnap=nsnap, fields=fields, grid_num=5,
                               maxs=(5, 5, 5, 5, 100000, 100000))
    n = tuple(cfg.n)
    S.run()
    st = S.stats()
    print("coarsened full channel: %d rays, %d segments, %.1f ms"
          % (st["n_rays"], st["n_steps"], st["ms_trace"]), flush=True)
    assert st["n_rays"] == n[0] * n[1] * n[2] * R * nsnap
    for s in range(nsnap):
        res = S.results(s)
        assert np.all(np.isfinite(res["Q"])) and np.all(res["E"] > 0)
        d = oracle_problem_dict(S, cfg, s)
        P = oracle.problem(d)
        for cell in _sample(n, 8, 200 + s):
            descs = oracle.gen_philox(P, cfg.seed, s, cell, 0, R)
            E, A, _, _ = oracle.trace_descs(P, cell, descs)
            assert abs(res["E"][cell] - E) <= 1e-5 * abs(E), (s, cell)
            assert abs(res["A"][cell] - A) <= 1e-5 * abs(A), (s, cell)
            assert abs(res["Q"][cell] - (E - A)) <= 1e-5 * abs(E), (s, cell)
    S.close()
