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


def declared_signatures():
    """{name: [C parameter types]} from the header's prototypes."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for ret, name, params in re.findall(r"([A-Za-z_][A-Za-z0-9_ ]*?\**)\s*\b(matpl_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", txt):
        ps = [p.strip() for p in params.split(",") if p.strip() and p.strip() != "void"]
        out[name] = (ret.strip(), ps)
    return out


def _ctype_of(c_decl: str):
    c = c_decl.replace("const", "").strip()
    if "*" in c:
        return ctypes.c_void_p
    base = c.rsplit(" ", 1)[0].strip() if " " in c else c
    return {"int": ctypes.c_int, "int64_t": ctypes.c_int64, "double": ctypes.c_double, "int32_t": ctypes.c_int}[base]


def test_header_matches_binding_table():
    assert declared_symbols() == sorted(_lib.exported_symbols())
    sigs = declared_signatures()
    assert sorted(sigs) == declared_symbols()
    for name, (ret, params) in sigs.items():
        restype, argtypes = _lib.SIGNATURES[name]
        want = [_ctype_of(p) for p in params]
        assert len(argtypes) == len(want), (name, len(argtypes), len(want))
        for k, (a, b) in enumerate(zip(argtypes, want)):
            assert a == b, (name, k, params[k])


def test_library_loads_and_exports_everything():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmatpl_hip.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert _lib.lib().matpl_abi_version() == 2


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libmatpl_hip.so")
    monkeypatch.setattr(_lib, "_LIB", None)
    with pytest.raises(_lib.MatplError):
        _lib.lib()
