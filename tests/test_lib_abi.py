"""CPU checks of the C-ABI library: it loads, and exports every declared symbol.

No compute calls are made (there is no GPU in the CPU test run).
"""
import ctypes
import os
import re

import pytest

from matpl_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "matpl_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(matpl_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_table():
    assert declared_symbols() == sorted(_lib.exported_symbols())


def test_library_loads_and_exports_everything():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmatpl_hip.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert _lib.lib().matpl_abi_version() == 1


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libmatpl_hip.so")
    monkeypatch.setattr(_lib, "_LIB", None)
    with pytest.raises(_lib.MatplError):
        _lib.lib()
