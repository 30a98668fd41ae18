"""The C ABI library (dnsmc_amd/libdnsmc_rad.so) on the CPU: it loads, exports
every function include/dnsmc_rad.h declares, its configuration defaults are
the reference's shipped code/param.h values, the key = value parser and its
errors behave, and the product never touches the oracle.  No compute call is
made here (that needs a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "dnsmc_rad.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set(re.findall(r"\b((?:dnsmc_\w+)|mc_gpu_?|get_results_?|mc_gpu_commit_|fix_temp_gpu_|fix_rad_gpu_)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import dnsmc_amd
    from dnsmc_amd._lib import EXPORTED, LIB_PATH
    L = ctypes.CDLL(LIB_PATH)
    names = _declared_functions()
    assert {"mc_gpu_", "get_results_", "mc_gpu", "get_results", "dnsmc_run"} <= set(names)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(EXPORTED) <= set(names)


def test_defaults_are_reference_param_h():
    import dnsmc_amd
    c = dnsmc_amd.default_config()
    assert list(c.n) == [64, 64, 64]                 # imax, jmax, kmax
    assert list(c.L) == [1.0, 1.0, 1.0]              # Lx, Ly, Lz
    assert c.nvar == 3 and c.photmax == 20000 and c.grid_num == 5
    assert list(c.maxs)[:5] == [5, 5, 5, 5, 100000]
    assert abs(c.toll - 1e-3) < 1e-9
    assert [list(r) for r in c.bc] == [[2, 2], [1, 1], [1, 1]]   # bdw bde / bds bdn / bdb bdt
    assert list(c.Tw[0]) == [500.0, 500.0]           # Tww, Twe
    assert c.num_streams == 16 and c.num_gpu == 1 and c.p_row == 1 and c.sobol == 1
    assert c.spectral == 2                            # grey = 2 (particle tables)


def test_parse_config_roundtrip_and_errors():
    import dnsmc_amd
    from dnsmc_amd._lib import DnsmcError
    c = dnsmc_amd.parse_config("n = 32 48 16\nphotmax = 128\nseed = 77  # comment\nbc = 1 1 2 2 1 1\n")
    assert list(c.n) == [32, 48, 16] and c.photmax == 128 and c.seed == 77
    assert [list(r) for r in c.bc] == [[1, 1], [2, 2], [1, 1]]
    with pytest.raises(DnsmcError):
        dnsmc_amd.parse_config("no_such_key = 3\n")
    with pytest.raises(DnsmcError):
        dnsmc_amd.parse_config("n = 32 x 16\n")


def test_create_without_device_fails_cleanly():
    from conftest import have_gpu
    if have_gpu():
        pytest.skip("a HIP device is present")
    import dnsmc_amd
    from dnsmc_amd._lib import DnsmcError
    with pytest.raises(DnsmcError) as e:
        dnsmc_amd.Solver(dnsmc_amd.default_config())
    assert e.value.code == -4                         # DNSMC_EDEVICE


def test_invalid_configs_are_rejected_before_device_use():
    import dnsmc_amd
    from dnsmc_amd._lib import DnsmcError
    for bad in ("n = 0 8 8\n", "toll = 1.5\n", "mode = 7\n", "n = 2000 8 8\n", "tally = 9\n"):
        c = dnsmc_amd.parse_config(bad)
        with pytest.raises(DnsmcError) as e:
            dnsmc_amd.Solver(c)
        assert e.value.code == -1, bad               # DNSMC_EINVAL


def test_product_never_uses_the_oracle():
    files = glob.glob(os.path.join(ROOT, "dnsmc_amd", "**", "*.py"), recursive=True)
    files += glob.glob(os.path.join(ROOT, "dnsmc_amd", "csrc", "*"))
    for f in files:
        text = open(f, errors="replace").read()
        assert "pyoracle" not in text and "liboracle" not in text and "oracle/" not in text, f


def test_bench_refuses_more_than_one_gpu():
    """bench.py is a one-GPU measurement: --gpus 2 exits with status 2 and a
    message before it touches the library or a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
