"""The C-ABI library loads without a GPU and exports every symbol the header declares."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "edgefem_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(efm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "efm_topology_create" in syms and "efm_solver_solve" in syms
    assert len(syms) >= 18


def test_library_exports_every_declared_symbol():
    from edgefem_amd import _capi
    if not os.path.exists(_capi.LIB_PATH):
        pytest.fail("libedgefem_amd.so is not built (python -m edgefem_amd.build)")
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_capi.EXPORTS) == set(declared_symbols())


def test_version_and_error_string_without_gpu():
    from edgefem_amd import _capi
    lib = _capi.load_library()
    assert lib.efm_version() >= 1
    assert isinstance(lib.efm_last_error(), bytes)
    # argument validation happens before any device call
    h = ctypes.c_void_p()
    rc = lib.efm_topology_create(0, None, 0, None, 0, None, None, ctypes.byref(h))
    assert rc == 1 and b"no nodes" in lib.efm_last_error()
