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
