"""GPU-resident DNS coupling: fix_temp / fix_rad slabs (code/fix_var.f:49-121)
through the C ABI against the coupling oracle."""
import numpy as np
import pytest

import dnsmc_amd
from dnsmc_amd import workloads
from oracle import coupling

pytestmark = pytest.mark.gpu

N = (16, 24, 12)
TC, TPLUS = 573.0, 1.0


def _slabs(pj, pk, seed=5):
    """x-pencil decomposition pj x pk: every rank holds all i."""
    rng = np.random.default_rng(seed)
    glob = rng.uniform(0.0, 0.6, size=(N[0] + 2, N[1], N[2]))
    out = []
    for jr in np.array_split(np.arange(1, N[1] + 1), pj):
        for kr in np.array_split(np.arange(1, N[2] + 1), pk):
            s = (1, int(jr[0]), int(kr[0]))
            e = (N[0], int(jr[-1]), int(kr[-1]))
            out.append((np.asfortranarray(glob[:, s[1] - 1:e[1], s[2] - 1:e[2]]), s, e))
    return out


def _solver(ck_table_dir, p_row=2):
    tdir, _ = ck_table_dir
    cfg = workloads.channel_config(N, rays_per_cell=64, spectral=0, Tw=573.0, table_dir=tdir, p_row=p_row)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    return S


def test_fix_temp_slabs_build_the_field(ck_table_dir):
    """Four ranks' slabs written on the device give the field the reference's
    all-reduce gives, bit for bit, and the same solve as dnsmc_set_field."""
    slabs = _slabs(2, 2)
    tmc = coupling.fix_temp(slabs, N, TC, TPLUS)
    S = _solver(ck_table_dir)
    S.set_emitters_slab(1)
    for c, s, e in slabs:
        S.fix_temp_slab(0, c, s, e, TC, TPLUS)
    S.commit_field(0)
    dev = S.array("T", 0).reshape(N[::-1])            # (k, j, i)
    np.testing.assert_array_equal(dev, np.transpose(tmc[1:-1, 1:-1, 1:-1], (2, 1, 0)))
    S.run()
    r_slab = S.results(0)
    S.set_field(0, tmc)
    S.run()
    r_set = S.results(0)
    np.testing.assert_allclose(r_slab["E"], r_set["E"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(r_slab["A"], r_set["A"], rtol=1e-12, atol=0)
    S.close()


def test_fix_temp_uncovered_cells_fail_the_commit(ck_table_dir):
    slabs = _slabs(2, 2)
    S = _solver(ck_table_dir)
    for c, s, e in slabs[:3]:
        S.fix_temp_slab(0, c, s, e, TC, TPLUS)
    with pytest.raises(dnsmc_amd.DnsmcError):
        S.commit_field(0)
    with pytest.raises(dnsmc_amd.DnsmcError):   # the field was not accepted
        S.run()
    for c, s, e in slabs:   # a failed commit ends the assembly: start again
        S.fix_temp_slab(0, c, s, e, TC, TPLUS)
    S.commit_field(0)
    S.close()


def test_fix_rad_slabs(ck_table_dir):
    """Each rank's divQ / vcore from the device results equals the reference's
    formula on the FP64 results (bit for bit), and on the real(4) copy the
    reference divides to FP32 rounding."""
    slabs = _slabs(2, 2)
    S = _solver(ck_table_dir, p_row=2)
    ystart = 13                                           # second node's j slab
    S.set_emitters_slab(ystart)
    S.set_field(0, coupling.fix_temp(slabs, N, TC, TPLUS))
    S.run()
    r = S.results(0)
    nj = N[1] // 2
    # emitters of the slab are numbered k, then j, then i (i fastest)
    q64 = np.zeros((N[0] + 2, nj + 2, N[2] + 2), order="F")
    v64 = np.zeros_like(q64)
    q64[1:-1, 1:-1, 1:-1] = np.transpose(r["Q"].reshape(N[2], nj, N[0]), (2, 1, 0))
    v64[1:-1, 1:-1, 1:-1] = np.transpose(r["sdev"].reshape(N[2], nj, N[0]), (2, 1, 0))
    q32, v32 = S.results_fortran(0)
    for k0, k1 in ((1, 5), (6, 12)):
        st, en = (1, ystart, k0), (N[0], ystart + nj - 1, k1)
        dq, vc = S.fix_rad_slab(0, st, en, TC)
        np.testing.assert_array_equal(dq, coupling.fix_rad(q64, st, en, TC, N[0], nj))
        np.testing.assert_array_equal(vc, coupling.fix_rad(v64, st, en, TC, N[0], nj))
        np.testing.assert_allclose(dq, coupling.fix_rad(q32, st, en, TC, N[0], nj), rtol=1e-7, atol=0)
    S.close()


def test_legacy_coupling_calls(tmp_path):
    """fix_temp_gpu_ per rank + mc_gpu_commit_ + fix_rad_gpu_ on the legacy
    solver against the reference's sequence fix_temp -> mc_gpu_ ->
    get_results_ -> fix_rad (code/main_comp.f:192-202) on the same solver."""
    import os
    from helpers import legacy_conf_text
    conf = tmp_path / "rad.conf"
    conf.write_text(legacy_conf_text())
    os.environ["DNSMC_CONFIG"] = str(conf)
    cfg = dnsmc_amd.parse_config(conf.read_text())
    n = tuple(cfg.n)
    nj = n[1] // cfg.p_row
    rng = np.random.default_rng(9)
    glob = rng.uniform(0.0, 0.5, size=(n[0] + 2, n[1], n[2]))
    Tc, Tplus = 600.0, 1.0
    slabs = []
    for jr in np.array_split(np.arange(1, n[1] + 1), 2):
        for kr in np.array_split(np.arange(1, n[2] + 1), 2):
            s = (1, int(jr[0]), int(kr[0]))
            e = (n[0], int(jr[-1]), int(kr[-1]))
            slabs.append((np.asfortranarray(glob[:, s[1] - 1:e[1], s[2] - 1:e[2]]), s, e))
    # reference sequence on the legacy solver
    tmc = coupling.fix_temp(slabs, n, Tc, Tplus)
    dnsmc_amd.mc_gpu(tmc, 1)
    Q = np.zeros((n[0] + 2, nj + 2, n[2] + 2), np.float32, order="F")
    V = np.zeros_like(Q)
    dnsmc_amd.get_results(Q, V)
    # the same step through the slab calls
    for c, s, e in slabs:
        dnsmc_amd.fix_temp_gpu(c, s, e, Tc, Tplus)
    dnsmc_amd.mc_gpu_commit(1)
    for c, s, e in slabs:
        dq, vc = dnsmc_amd.fix_rad_gpu(s, e, Tc, (n[0] + 2, nj, e[2] - s[2] + 1))
        np.testing.assert_allclose(dq, coupling.fix_rad(Q, s, e, Tc, n[0], nj), rtol=1e-7, atol=0)
        np.testing.assert_allclose(vc, coupling.fix_rad(V, s, e, Tc, n[0], nj), rtol=1e-7, atol=0)
        assert np.count_nonzero(dq) == n[0] * nj * (e[2] - s[2] + 1)
