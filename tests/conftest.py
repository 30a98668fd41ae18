import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _hip_device_count():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = ctypes.CDLL(name)
        except OSError:
            continue
        n = ctypes.c_int(0)
        try:
            if h.hipGetDeviceCount(ctypes.byref(n)) == 0:
                return n.value
        except Exception:
            return 0
    return 0


_NDEV = None


def have_gpu():
    global _NDEV
    if _NDEV is None:
        _NDEV = _hip_device_count() if os.path.exists("/dev/kfd") else 0
    return _NDEV > 0


def pytest_collection_modifyitems(config, items):
    for it in items:
        if "gpu" in it.keywords and not have_gpu():
            it.add_marker(pytest.mark.skip(reason="no HIP device"))


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.build()
    return pyoracle.get()


@pytest.fixture(scope="session")
def ck_table_dir(tmp_path_factory):
    """Synthetic 100x16x69 correlated-k table written in the reference formats."""
    from dnsmc_amd import workloads
    d = tmp_path_factory.mktemp("ck_tables")
    t = workloads.synth_ck_table(11)
    workloads.write_reference_tables(t, str(d))
    return str(d), t
