"""The C-ABI library loads and exports every symbol the header declares (no GPU needed)."""

import ctypes
import os
import re

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
