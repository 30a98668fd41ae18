"""dnsmc_run_noalloc, the allocation-free variant of dnsmc_run (include/dnsmc_rad.h):
called through the C ABI with caller-owned device output arrays, its results
must equal dnsmc_run's on the same inputs, on two calls in a row (the second
with fewer emitters, so the shared scratch array still holds the first call's
tallies), and it must refuse what it cannot do without allocating."""
import ctypes as C

import numpy as np
import pytest

import dnsmc_amd
from dnsmc_amd import workloads
from dnsmc_amd._lib import DnsmcError, check, lib
from dnsmc_amd.devarray import DeviceArray

pytestmark = pytest.mark.gpu

EINVAL, ESTATE = -1, -5
N = (16, 24, 16)


def _noalloc(S, nemit, nsnap):
    Q = DeviceArray((nsnap, nemit), np.float64)
    V = DeviceArray((nsnap, nemit), np.float64)
    check(lib().dnsmc_run_noalloc(S.h, C.c_void_p(Q.ptr), C.c_void_p(V.ptr)))
    S.wait()
    return Q.numpy(), V.numpy()


def _solver(tdir, nsnap=2):
    cfg = workloads.channel_config(N, rays_per_cell=64, spectral=0, Tw=573.0, table_dir=tdir, nsnap=nsnap)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    for sn in range(nsnap):
        S.set_field(sn, workloads.ghosted(workloads.field_profile_noise(5 + sn, N)))
    return S


def _same(Qa, Va, ref, steps_a, steps_ref):
    for sn, r in enumerate(ref):
        scale = np.max(np.abs(r["Q"]))
        assert np.max(np.abs(Qa[sn] - r["Q"])) <= 1e-9 * scale, sn
        np.testing.assert_allclose(Va[sn], r["sdev"], rtol=1e-6, atol=1e-9 * scale)
    assert steps_a == steps_ref


def test_noalloc_matches_run_twice(ck_table_dir):
    S = _solver(ck_table_dir[0])
    ncell = N[0] * N[1] * N[2]
    # the existing entry point sizes every buffer the allocation-free call reads
    S.run()
    ref = [S.results(sn) for sn in range(2)]
    steps_ref = S.stats()["n_steps"]
    Q, V = _noalloc(S, ncell, 2)
    _same(Q, V, ref, S.stats()["n_steps"], steps_ref)
    # second call with a smaller input: 1000 of the cells (the scratch array
    # still holds the first call's tallies; the slot buffers hold the subset)
    cells = np.sort(np.random.default_rng(3).choice(ncell, 1000, replace=False)).astype(np.int32)
    S.set_emitters(cells)
    Q2, V2 = _noalloc(S, len(cells), 2)
    steps2 = S.stats()["n_steps"]
    S.run()
    ref2 = [S.results(sn) for sn in range(2)]
    _same(Q2, V2, ref2, steps2, S.stats()["n_steps"])
    # the subset's results are those cells' results of the full solve
    for sn in range(2):
        np.testing.assert_allclose(Q2[sn], ref[sn]["Q"][cells], rtol=1e-9, atol=1e-9 * np.max(np.abs(ref[sn]["Q"])))
    S.close()


def test_noalloc_split_launches(ck_table_dir, monkeypatch):
    """A solve cut into many traversal launches (97 slots each, the split a
    snapshot of more than 2^32 rays takes): consecutive launches alternate over
    two streams with the claim counters at two places of the scratch array;
    the results equal dnsmc_run's, split the same way and whole."""
    S = _solver(ck_table_dir[0])
    ncell = N[0] * N[1] * N[2]
    S.run()
    whole = [S.results(sn) for sn in range(2)]
    steps_whole = S.stats()["n_steps"]
    monkeypatch.setenv("DNSMC_FUSED_SLOTS", "97")
    S.run()
    ref = [S.results(sn) for sn in range(2)]
    st = S.stats()
    assert st["n_trace_launches"] == -(-ncell // 97)
    _same([r["Q"] for r in ref], [r["sdev"] for r in ref], whole, st["n_steps"], steps_whole)
    Q, V = _noalloc(S, ncell, 2)
    _same(Q, V, ref, S.stats()["n_steps"], st["n_steps"])
    S.close()


def test_noalloc_refusals(ck_table_dir):
    # nothing sized yet: refused with the buffer named, the solver still usable
    S = _solver(ck_table_dir[0], nsnap=1)
    with pytest.raises(DnsmcError) as e:
        _noalloc(S, N[0] * N[1] * N[2], 1)
    assert e.value.code == ESTATE and "would have to grow" in str(e.value)
    S.run()
    assert np.all(np.isfinite(S.results(0)["Q"]))
    S.close()
    # other estimators allocate descriptor buffers: refused
    cfg = workloads.channel_config(N, rays_per_cell=16, spectral=3, Tw=300.0, tally=1)
    S = dnsmc_amd.Solver(cfg)
    with pytest.raises(DnsmcError) as e:
        _noalloc(S, N[0] * N[1] * N[2], 1)
    assert e.value.code == EINVAL
    S.close()
    # beyond the scratch capacity (2,796,032 emitters x snapshots): refused
    big = (128, 128, 96)
    cfg = workloads.channel_config(big, rays_per_cell=16, spectral=3, Tw=300.0, nsnap=2)
    S = dnsmc_amd.Solver(cfg)
    with pytest.raises(DnsmcError) as e:
        _noalloc(S, big[0] * big[1] * big[2], 2)
    assert e.value.code == EINVAL and "scratch" in str(e.value)
    S.close()
