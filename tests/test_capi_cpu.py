"""The C-ABI library loads and exports every symbol the header declares (no GPU needed)."""

import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "mumaxplus_amd.h")
LIB = os.path.join(ROOT, "mumax-plus_amd", "mumaxplus", "libmumaxplus_amd.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mpa_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "mpa_ts_steps" in names and "mpa_demag_plan_exec" in names
    assert len(names) > 40


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    import mumaxplus._mumaxpluscpp as b
    assert set(declared()) == set(b.EXPORTED)


def test_error_path_without_device():
    """A call with a NULL handle reports an error instead of crashing."""
    import mumaxplus._mumaxpluscpp as b
    rc = b._lib.mpa_ts_step(None)
    assert rc == b.MPA_ERR_INVALID
    assert b"NULL" in b._lib.mpa_last_error()


def test_batch_tuning_record_round_trip(tmp_path):
    """The step-path autotuner's record: a file merged into the (process-wide)
    record and written back gives the same entries (no device needed)."""
    from mumaxplus import _mumaxpluscpp as cpp

    src = tmp_path / "tune.txt"
    src.write_text("cells<2^12 members<2^2 line<=2^6 3d DormandPrince\t0\t1 2 0.19\t1 2 0.73\t1 2 0.41\t1 2 0.31\n"
                   "cells<2^21 members<2^9 line<=2^7 3d DormandPrince\t-1\t1 1 1.3\t0 0 0\t0 0 0\t0 0 0\n")
    cpp._batch_tune_clear()
    cpp._batch_tune_load(src)
    out = tmp_path / "out.txt"
    cpp._batch_tune_save(out)
    rows = sorted(line.split("\t") for line in out.read_text().splitlines())
    want = sorted(line.split("\t") for line in src.read_text().splitlines())
    assert [r[0] for r in rows] == [w[0] for w in want]
    for r, w in zip(rows, want):
        assert int(r[1]) == int(w[1])
        for a, b in zip(r[2:], w[2:]):
            ua, ca, ba = a.split()
            ub, cb, bb = b.split()
            assert (ua, ca) == (ub, cb) and float(ba) == float(bb)
    cpp._batch_tune_clear()
    bad = tmp_path / "bad.txt"
    bad.write_text("no tab here\n")
    with pytest.raises(ValueError, match="malformed"):
        cpp._batch_tune_load(bad)
    with pytest.raises(RuntimeError, match="cannot read"):
        cpp._batch_tune_load(tmp_path / "missing.txt")
