"""ctypes access to oracle/_ref/libnepref.so (TEST INFRASTRUCTURE).

The library is the reference's own NEP CPU code compiled from its sources by
oracle/ref.mk (src/op/src/nep_cpu.cpp, src/op/src/cpu_calculate_nepneighbor.cpp).
It is used only to check the oracle's NEP restatement; ``available()`` is
False where it has not been built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(_HERE, "_ref", "libnepref.so")
_L = None

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_d = ctypes.c_double
_l = ctypes.c_int64

SYM = {
    "neighbor": "_Z29launch_calculate_neighbor_cpuPKlS0_S0_S0_PKdS2_S0_S2_ddllllPlS3_S3_S3_PdS4_b",
    "descriptor": "_Z31launch_calculate_descriptor_cpuPKdS0_S0_PKlS2_ddPdllllllllll",
    "maxneigh": "_Z29launch_calculate_maxneigh_cpuPKlS0_PKdS2_S0_S2_ddllPlS3_lbS0_",
    "find_fc": "_Z7find_fcdddRd",
    "find_fn": "_Z7find_fnidddPd",
    "accumulate_s": "_Z12accumulate_sdddddPd",
    "find_q": "_Z6find_qiiPKdPd",
    "find_q_with_4body": "_Z17find_q_with_4bodyiiPKdPd",
    "find_q_with_5body": "_Z17find_q_with_5bodyiiPKdPd",
}


def available() -> bool:
    return os.path.exists(PATH)


def lib():
    global _L
    if _L is None:
        L = ctypes.CDLL(PATH)
        f = getattr(L, SYM["neighbor"])
        f.argtypes = [_i64p, _i64p, _i64p, _i64p, _f64p, _f64p, _i64p, _f64p, _d, _d, _l, _l, _l, _l, _i64p, _i64p,
                      _i64p, _i64p, _f64p, _f64p, ctypes.c_bool]
        f = getattr(L, SYM["descriptor"])
        f.argtypes = [_f64p, _f64p, _f64p, _i64p, _i64p, _d, _d, _f64p] + [_l] * 10
        f = getattr(L, SYM["maxneigh"])
        f.argtypes = [_i64p, _i64p, _f64p, _f64p, _i64p, _f64p, _d, _d, _l, _l, _i64p, _i64p, _l, ctypes.c_bool,
                      _i64p]
        getattr(L, SYM["find_fc"]).argtypes = [_d, _d, _d, ctypes.POINTER(_d)]
        getattr(L, SYM["find_fn"]).argtypes = [ctypes.c_int, _d, _d, _d, _f64p]
        getattr(L, SYM["accumulate_s"]).argtypes = [_d, _d, _d, _d, _d, _f64p]
        for k in ("find_q", "find_q_with_4body", "find_q_with_5body"):
            getattr(L, SYM[k]).argtypes = [ctypes.c_int, ctypes.c_int, _f64p, _f64p]
        _L = L
    return _L


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def neighbor(packed, rc_r, rc_a, maxr, maxa, with_rij=True):
    n = int(packed["num_atoms"].sum())
    w = 4 if with_rij else 3
    out = dict(NN_r=np.zeros(n, np.int64), NL_r=np.full((n, maxr), -1, np.int64), NN_a=np.zeros(n, np.int64),
               NL_a=np.full((n, maxa), -1, np.int64), Ri_r=np.zeros((n, maxr, w)), Ri_a=np.zeros((n, maxa, w)))
    nums = _c(packed["num_atoms"], np.int64)
    getattr(lib(), SYM["neighbor"])(nums, _c(np.cumsum(nums), np.int64), _c(packed["atom_type_map"], np.int64),
                                    _c(packed["atom_type_map"], np.int64), _c(packed["box"], np.float64),
                                    _c(packed["box_orig"], np.float64), _c(packed["num_cell"], np.int64),
                                    _c(packed["position"], np.float64), float(rc_r), float(rc_a), maxr, maxa,
                                    len(nums), n, out["NN_r"], out["NL_r"], out["NN_a"], out["NL_a"], out["Ri_r"],
                                    out["Ri_a"], bool(with_rij))
    return out


def maxneigh(packed, rc_r, rc_a, ntypes, with_type):
    n = int(packed["num_atoms"].sum())
    shape = (n, ntypes) if with_type else (n,)
    NN_r = np.zeros(shape, np.int64)
    NN_a = np.zeros(shape, np.int64)
    nums = _c(packed["num_atoms"], np.int64)
    getattr(lib(), SYM["maxneigh"])(nums, _c(np.cumsum(nums), np.int64), _c(packed["box"], np.float64),
                                    _c(packed["box_orig"], np.float64), _c(packed["num_cell"], np.int64),
                                    _c(packed["position"], np.float64), float(rc_r), float(rc_a), len(nums), n, NN_r,
                                    NN_a, ntypes, bool(with_type), _c(packed["atom_type_map"], np.int64))
    return NN_r, NN_a


def descriptor(c2, c3, rij_xyz, NL, atom_map, rc_r, rc_a, lmax3, lmax4, lmax5):
    """calculate_descriptor: radial + angular features from the radial list and [x,y,z]."""
    c2 = _c(c2, np.float64)
    c3 = _c(c3, np.float64)
    T, _, nm2, nb2 = c2.shape
    nm3, nb3 = c3.shape[2], c3.shape[3]
    n, S = NL.shape
    f3 = (nm3 * lmax3 if lmax3 > 0 else 0) + (nm3 if lmax4 > 0 else 0) + (nm3 if lmax5 > 0 else 0)
    feats = np.zeros((n, nm2 + f3))
    getattr(lib(), SYM["descriptor"])(c2, c3, _c(rij_xyz, np.float64), _c(NL, np.int64), _c(atom_map, np.int64),
                                      float(rc_r), float(rc_a), feats, n, S, nm2, nb2, nm3, nb3, lmax3, lmax4, lmax5,
                                      T)
    return feats


def find_fc(rc, d12):
    out = _d(0.0)
    getattr(lib(), SYM["find_fc"])(rc, 1.0 / rc, d12, ctypes.byref(out))
    return out.value


def find_fn(n_max, rc, d12, fc):
    fn = np.zeros(n_max + 1)
    getattr(lib(), SYM["find_fn"])(n_max, 1.0 / rc, d12, fc, fn)
    return fn


def accumulate_s(d12, x, y, z, fn):
    s = np.zeros(24)
    getattr(lib(), SYM["accumulate_s"])(d12, x, y, z, fn, s)
    return s


def find_q(kind, na, n, s):
    q = np.zeros(6 * na)
    getattr(lib(), SYM[kind])(na, n, _c(s, np.float64), q)
    return q
