"""Error behaviour of the extended C ABI (include/dnsmc_rad.h): every misuse
returns a DNSMC_E* code with a message and leaves the solver usable; the
conditions are forced from the host, no kernel is made to fail."""
import numpy as np
import pytest

import dnsmc_amd
from dnsmc_amd import workloads

pytestmark = pytest.mark.gpu

EINVAL, EIO, ERANGE, EDEVICE, ESTATE = -1, -2, -3, -4, -5
N = (8, 12, 8)


def _solver(**kw):
    cfg = workloads.channel_config(N, rays_per_cell=16, spectral=3, Tw=300.0, **kw)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    return S


def _code(fn, *a):
    with pytest.raises(dnsmc_amd.DnsmcError) as e:
        fn(*a)
    return e.value.code


def _field(T=1000.0):
    return workloads.ghosted(np.full(N, T, np.float32))


def test_state_errors_then_recovery():
    S = _solver()
    assert _code(S.run) == ESTATE                  # no temperature field yet
    assert _code(S.results, 0) == ESTATE           # no solve yet
    assert _code(S.fix_rad_slab, 0, (1, 1, 1), (N[0], N[1], 2), 573.0) == ESTATE
    assert _code(S.commit_field, 0) == ESTATE      # no slab written
    S.set_field(0, _field())
    S.run()
    r = S.results(0)
    assert np.all(np.isfinite(r["Q"]))
    S.close()


def test_field_errors():
    S = _solver()
    bad = _field()
    bad[3, 4, 5] = np.nan
    assert _code(S.set_field, 0, bad) == ERANGE
    bad = _field()
    bad[2, 2, 2] = -5.0
    assert _code(S.set_field, 0, bad) == ERANGE
    assert _code(S.set_field, 1, _field()) == EINVAL   # one snapshot configured
    # slabs: a non-positive temperature is caught at the commit
    c = np.full((N[0] + 2, N[1], N[2]), -2.0)          # T = c * Tc / Tplus + Tc < 0
    S.fix_temp_slab(0, c, (1, 1, 1), (N[0], N[1], N[2]), 300.0, 1.0)
    assert _code(S.commit_field, 0) == ERANGE
    assert _code(S.fix_temp_slab, 0, c, (0, 1, 1), (N[0], N[1], N[2]), 300.0, 1.0) == EINVAL
    assert _code(S.fix_temp_slab, 0, c, (1, 1, 1), (N[0], N[1], N[2]), 300.0, 0.0) == EINVAL
    S.set_field(0, _field())                           # still usable
    S.run()
    S.close()


def test_emitter_and_mode_errors():
    S = _solver()
    assert _code(S.set_emitters, np.array([N[0] * N[1] * N[2]])) == EINVAL
    assert _code(S.set_emitters, np.array([-1])) == EINVAL
    assert _code(S.set_emitters_slab, N[1] + 1) == EINVAL
    S.close()
    L = _solver(mode=2)
    L.set_field(0, _field())
    assert _code(L.run) == ESTATE                      # learned mode without a network
    L.close()


def test_table_errors_through_the_solver(ck_table_dir, tmp_path):
    """A missing table directory is DNSMC_EIO at load time (the reference's
    readT opens its files unchecked, code/read_tables.cpp:25,44)."""
    cfg = workloads.channel_config(N, rays_per_cell=16, spectral=0, Tw=573.0, table_dir=str(tmp_path / "none"))
    S = dnsmc_amd.Solver(cfg)
    assert _code(S.load_tables) == EIO
    S.close()


def test_tmax_outside_the_table():
    """Tmax at or above the table's last node (2000 K) is DNSMC_ERANGE at run
    time, from the field or from a wall (Tmax includes the wall temperatures,
    code/mc_gpu.cu:481-486); the reference would read past its proposal rows
    (code/mc_gpu.cu:584-589)."""
    from dnsmc_amd import workloads as w
    t = w.synth_ck_table(11)
    cfg = w.channel_config(N, rays_per_cell=16, spectral=0, Tw=573.0)
    S = dnsmc_amd.Solver(cfg)
    S.set_tables(w.table_arrays(t))
    hot = _field(1000.0)
    hot[4, 5, 6] = 2000.0
    S.set_field(0, hot)
    assert _code(S.run) == ERANGE
    S.set_field(0, _field(1999.0))
    S.run()                                            # just below the last node: fine
    assert np.all(np.isfinite(S.results(0)["Q"]))
    S.close()
    cfg = w.channel_config(N, rays_per_cell=16, spectral=0, Tw=2050.0)
    S = dnsmc_amd.Solver(cfg)
    S.set_tables(w.table_arrays(t))
    S.set_field(0, _field(1000.0))
    assert _code(S.run) == ERANGE                      # the wall is above the table
    S.close()


def test_slot_count_limit():
    """Tally slots are 32-bit: emitters x blocks x snapshots >= 2^31 is refused
    before anything is allocated (compat mode, 64^3 cells x 8192 blocks)."""
    n = (64, 64, 64)
    cfg = workloads.channel_config(n, spectral=3, Tw=300.0, mode=1, nvar=8192, photmax=1)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    S.set_field(0, workloads.ghosted(np.full(n, 1000.0, np.float32)))
    assert _code(S.run) == EINVAL
    S.close()


def test_results_keep_the_slab_of_their_run():
    """Changing the emitter slab after a run does not move that run's results:
    the Fortran-layout and coupling outputs place them by the run's ystart."""
    cfg = workloads.channel_config(N, rays_per_cell=16, spectral=3, Tw=300.0, p_row=2)
    S = dnsmc_amd.Solver(cfg)
    S.load_tables()
    x = np.linspace(900.0, 1100.0, N[1])
    S.set_field(0, workloads.ghosted(np.broadcast_to(x[None, :, None], N).astype(np.float32)))
    S.set_emitters_slab(7)
    S.run()
    Q1, V1 = S.results_fortran(0)
    d1, v1 = S.fix_rad_slab(0, (1, 7, 1), (N[0], N[1], N[2]), 573.0)
    S.set_emitters_slab(1)                             # no run: results stay those of ystart = 7
    Q2, V2 = S.results_fortran(0)
    d2, v2 = S.fix_rad_slab(0, (1, 7, 1), (N[0], N[1], N[2]), 573.0)
    np.testing.assert_array_equal(Q1, Q2)
    np.testing.assert_array_equal(V1, V2)
    np.testing.assert_array_equal(d1, d2)
    np.testing.assert_array_equal(v1, v2)
    assert np.count_nonzero(d1) == N[0] * (N[1] // 2) * N[2]
    S.close()
