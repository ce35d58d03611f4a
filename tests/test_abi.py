"""The C-ABI library loads without a GPU and exports every entry point that
include/tetmag_hip.h declares (no compute calls are made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tetmag_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(tm_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert len(names) > 50
    for must in ("tm_ctx_create", "tm_eval_rhs", "tm_bem_apply", "tm_solver_solve", "tm_batch_eval_rhs"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from tetmag_amd import _native as N
    lib = N.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_signatures_cover_header():
    from tetmag_amd import _native as N
    missing = [s for s in declared_symbols() if s not in N._SIGNATURES]
    assert not missing, missing


def test_version_and_no_device_paths():
    from tetmag_amd import _native as N
    lib = N.lib()
    assert lib.tm_version().decode().startswith("tetmag_hip")
    with pytest.raises(N.TetmagError):
        N.check(N.TM_ERR_BAD_ARG, "probe")
