"""GPU tests of the achievable-bandwidth hook (dnsmc_bench_stream): the read
and copy ceilings are positive and below the MI355X's 8 TB/s HBM peak over
arrays larger than its 256 MB last-level cache, the winning variant is one of
the twelve, and bad arguments are refused before any device work."""
import ctypes as C

import pytest

from dnsmc_amd._lib import DNSMC_EINVAL, lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", [0, 1])
def test_stream_ceiling_is_plausible(kind):
    g = C.c_double(0.0)
    best = C.c_int(-1)
    assert lib().dnsmc_bench_stream(2 << 30, 3, kind, C.byref(g), C.byref(best)) == 0
    assert 1000.0 < g.value < 8000.0 * 1.02, g.value
    assert 0 <= best.value < 12


def test_stream_refuses_bad_arguments():
    g = C.c_double(0.0)
    for nbytes, reps, kind in ((8, 1, 0), (1 << 20, 0, 0), (1 << 20, 1, 2), (1 << 20, 1, -1)):
        assert lib().dnsmc_bench_stream(nbytes, reps, kind, C.byref(g), None) == DNSMC_EINVAL
