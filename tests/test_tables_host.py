"""Host-side table reading (no device): the library's reader of the reference
table formats (code/read_tables.cpp:20-92 gas, :187-241 particle, :104-183
grey==1 synthesis) through dnsmc_probe_tables.

The reference opens and scans the files without checking (code/read_tables.cpp:25,
:44, :70); the library reports a missing, short, malformed or inconsistent file
as DNSMC_EIO with the file named in the message.  The files are the
reference's own particle tables (tests/golden/ref_particle_tables) and the
synthetic 100x16x69 correlated-k table written in the reference formats.
"""
import os
import shutil

import pytest

import dnsmc_amd
from dnsmc_amd import workloads
from helpers import particle_table_dir

EIO = -2


def _cfg(spectral, tdir):
    return workloads.channel_config((8, 12, 8), spectral=spectral, table_dir=tdir)


def _eio(spectral, tdir):
    with pytest.raises(dnsmc_amd.DnsmcError) as e:
        dnsmc_amd.probe_tables(_cfg(spectral, tdir))
    assert e.value.code == EIO, str(e.value)
    return str(e.value)


def _copy(src, tmp_path, name):
    d = tmp_path / name
    shutil.copytree(src, d)
    return d


def test_reads_reference_and_synthetic_tables(ck_table_dir):
    assert dnsmc_amd.probe_tables(_cfg(2, particle_table_dir())) == (119, 1, 52)
    assert dnsmc_amd.probe_tables(_cfg(0, ck_table_dir[0])) == (100, 16, 69)
    # grey==1 tables are synthesised at run time: 1000 bands, 55 temperatures
    assert dnsmc_amd.probe_tables(_cfg(1, "/nonexistent")) == (1000, 1, 55)


def test_missing_directory_and_files(ck_table_dir, tmp_path):
    assert "planck-mean.txt" in _eio(0, str(tmp_path / "nowhere"))
    d = _copy(ck_table_dir[0], tmp_path, "no_prob")
    os.remove(d / "prob_new2.txt")
    assert "prob_new2.txt" in _eio(0, str(d))
    d = _copy(particle_table_dir(), tmp_path, "no_pprob")
    os.remove(d / "prob.txt")
    assert "prob.txt" in _eio(2, str(d))
    d = _copy(ck_table_dir[0], tmp_path, "no_bands")
    for f in os.listdir(d):
        if f.startswith("NarrBand"):
            os.remove(d / f)
    assert "NarrBand" in _eio(0, str(d))


def test_band_file_missing_from_the_sequence(ck_table_dir, tmp_path):
    """NarrBand5.txt absent: the band count stops at 5 while the probability
    file still holds 100 bands per temperature (the reference would read a
    5-band table silently)."""
    d = _copy(ck_table_dir[0], tmp_path, "gap")
    os.remove(d / "NarrBand5.txt")
    msg = _eio(0, str(d))
    assert "prob_new2.txt" in msg


def _truncate(path, keep_lines):
    with open(path) as f:
        lines = f.readlines()
    with open(path, "w") as f:
        f.writelines(lines[:keep_lines])


def test_short_files(ck_table_dir, tmp_path):
    d = _copy(ck_table_dir[0], tmp_path, "short_band")
    _truncate(d / "NarrBand17.txt", 40)            # header + 39 of 69 rows
    assert "NarrBand17.txt" in _eio(0, str(d))
    d = _copy(ck_table_dir[0], tmp_path, "short_prob")
    _truncate(d / "prob_new2.txt", 69 * 100 - 3)
    assert "prob_new2.txt" in _eio(0, str(d))
    d = _copy(particle_table_dir(), tmp_path, "short_pprob")
    _truncate(d / "prob.txt", 52 * 119 - 1)
    assert "prob.txt" in _eio(2, str(d))


def test_malformed_and_inconsistent_rows(ck_table_dir, tmp_path):
    d = _copy(ck_table_dir[0], tmp_path, "bad_row")
    with open(d / "NarrBand3.txt") as f:
        lines = f.readlines()
    lines[10] = "garbage\n"
    with open(d / "NarrBand3.txt", "w") as f:
        f.writelines(lines)
    assert "NarrBand3.txt" in _eio(0, str(d))
    d = _copy(ck_table_dir[0], tmp_path, "nq")
    with open(d / "NarrBand8.txt") as f:
        lines = f.readlines()
    lines[5] = " ".join(lines[5].split()[:-2]) + "\n"   # 14 instead of 16 g columns
    with open(d / "NarrBand8.txt", "w") as f:
        f.writelines(lines)
    assert "NarrBand8.txt" in _eio(0, str(d))
    d = _copy(ck_table_dir[0], tmp_path, "long_prob")
    with open(d / "prob_new2.txt") as f:
        last = f.readlines()[-1]
    with open(d / "prob_new2.txt", "a") as f:
        f.write(last)
    assert "more rows" in _eio(0, str(d))
