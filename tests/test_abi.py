"""The C-ABI library loads and exports every symbol include/*.h declares, and
argument errors are reported as status codes (no GPU needed)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

import cosmos_amd
from cosmos_amd import native


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names.update(re.findall(r"\b(cosmos_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "cosmos_sweep" in names and "cosmos_solve" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = native.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) <= set(native.SIGNATURES)


def test_argument_errors_are_status_codes():
    lib = native.load()
    h = ctypes.c_void_p()
    assert lib.cosmos_create(ctypes.byref(h), 7, 2, 8, 8, 1.4) == -1
    assert b"physics" in lib.cosmos_last_error(None)
    assert lib.cosmos_create(ctypes.byref(h), 0, 3, 8, 8, 1.4) == -1
    assert lib.cosmos_create(ctypes.byref(h), 0, 2, 0, 8, 1.4) == -1
    assert lib.cosmos_create(None, 0, 2, 8, 8, 1.4) == -1
    assert lib.cosmos_sweep(None, 2, 0, 0.1, 0.1, 0.0, 1) == -1
    assert lib.cosmos_destroy(None) == 0


def test_python_binding_raises_on_error():
    with pytest.raises(cosmos_amd.CosmosError):
        cosmos_amd.Solver(9, 2, 4, 4)
