"""FP64 CPU oracle for the descriptor-and-force path (TEST INFRASTRUCTURE).

This package is the checker that the HIP path is compared against. Only the
test suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; the product package never does.

The C restatement lives in ``dp_oracle.c`` (and ``nep_oracle.c``); each C
routine cites the reference file:line it follows. This module only marshals
numpy arrays through ctypes and chains the routines into a whole DP
evaluation (energy, forces, virial) the way ``DP.forward`` does on the CPU
(src/model/dp_dp.py:167-215 with the CPU branch of :436-459).

Parity pins (tests/test_oracle_dp.py): fcc shell counts, brute-force
neighbor symmetry, central finite differences of Ri/forces/virial,
translation/rotation invariance, sum of forces = 0.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return path


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        c_int, c_long, c_double = ctypes.c_int, ctypes.c_long, ctypes.c_double
        L.oracle_dp_find_neighbor.argtypes = [_f64p, _f64p, _i32p, c_int, c_int, c_int, c_double,
                                              _f64p, _i32p, _f64p]
        L.oracle_dp_find_neighbor.restype = c_int
        L.oracle_dp_image_dr.argtypes = [_i32p, _f64p, c_long, _f64p]
        L.oracle_dp_env_mat.argtypes = [_f64p, _i32p, _f64p, _f64p, c_int, c_int, c_double, c_double,
                                        _f64p, _f64p]
        L.oracle_dp_calc_stat.argtypes = [_i32p, _f64p, _i32p, c_int, c_int, c_int, c_double, c_double,
                                          _f64p, _f64p]
        L.oracle_dp_energy_grad.argtypes = [_f64p, _i32p, _i32p, c_int, c_int, c_int, c_int,
                                            _i32p, _f64p, _i32p, _f64p, _f64p, _f64p]
        L.oracle_dp_energy_grad.restype = c_int
        L.oracle_dp_force_virial.argtypes = [_i32p, _f64p, _f64p, _f64p, c_int, c_int, c_int, _f64p, _f64p]
        L.oracle_dp_force_op.argtypes = [_i32p, _f64p, _f64p, _f64p, c_int, c_int, c_int, c_int, _f64p]
        L.oracle_dp_virial_op.argtypes = [_i32p, _f64p, _f64p, _f64p, c_int, c_int, c_int, c_int, _f64p, _f64p]
        L.oracle_dp_force_grad_op.argtypes = [_i32p, _f64p, _f64p, c_int, c_int, c_int, c_int, _f64p]
        L.oracle_dp_virial_grad_op.argtypes = [_i32p, _f64p, _f64p, _f64p, c_int, c_int, c_int, _f64p]
        L.oracle_num_threads.restype = c_int
        L.oracle_set_num_threads.argtypes = [c_int]
        L.oracle_set_num_threads.restype = None
        _LIB = L
    return _LIB


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# --------------------------------------------------------------------------- D1
def dp_find_neighbor(lattice, frac, type_map, ntypes, m, rc_m, rc_type=None):
    """NeighConst.find_neighbore for one image. Returns (list [N,S] int32 1-based, dR [N,S,3])."""
    lattice = _c(lattice, np.float64).reshape(3, 3)
    frac = _c(frac, np.float64).copy()
    type_map = _c(type_map, np.int32)
    n = frac.shape[0]
    if rc_type is None:
        rc_type = np.full(ntypes, rc_m)
    rc_type = _c(rc_type, np.float64)
    lst = np.zeros((n, ntypes * m), np.int32)
    dR = np.zeros((n, ntypes * m, 3), np.float64)
    err = lib().oracle_dp_find_neighbor(lattice, frac, type_map, ntypes, n, m, rc_m, rc_type, lst, dR)
    if err:
        raise OverflowError(f"maxNeighborNum {m} too small (atom {err - 1})")
    return lst, dR


def dp_image_dr(lst, dR):
    lst = _c(lst, np.int32)
    dR = _c(dR, np.float64)
    out = np.zeros(lst.shape + (4,), np.float64)
    lib().oracle_dp_image_dr(lst, dR, lst.size, out)
    return out


# --------------------------------------------------------------------------- D3/D4
def dp_env_mat(image_dr, type_map, davg, dstd, rmin, rmax):
    """Normalized Ri [N,S,4] and Ri_d [N,S,4,3]; davg/dstd [T,4] or [T,S*4]."""
    image_dr = _c(image_dr, np.float64)
    n, S = image_dr.shape[:2]
    T = davg.shape[0]
    davg = np.asarray(davg, np.float64)
    dstd = np.asarray(dstd, np.float64)
    if davg.shape[1] == 4:
        davg = np.tile(davg[:, None, :], (1, S, 1))
        dstd = np.tile(dstd[:, None, :], (1, S, 1))
    davg = _c(davg, np.float64).reshape(T, S * 4)
    dstd = _c(dstd, np.float64).reshape(T, S * 4)
    Ri = np.zeros((n, S, 4))
    Ri_d = np.zeros((n, S, 4, 3))
    lib().oracle_dp_env_mat(image_dr, _c(type_map, np.int32), davg, dstd, n, S, rmin, rmax, Ri, Ri_d)
    return Ri, Ri_d


def dp_calc_stat(lst, dR, type_map, ntypes, rmin, rmax):
    lst = _c(lst, np.int32)
    n, S = lst.shape
    davg = np.zeros((ntypes, 4))
    dstd = np.zeros((ntypes, 4))
    lib().oracle_dp_calc_stat(lst, _c(dR, np.float64), _c(type_map, np.int32), n, S, ntypes,
                              rmin, rmax, davg, dstd)
    return davg, dstd


# --------------------------------------------------------------------------- D5/D6
def _pack_nets(nets, sizes):
    desc = [len(sizes) - 1] + list(sizes)
    has_bias = any(l[1] is not None for net in nets for l in net)
    has_rdt = any(l[2] is not None for net in nets for l in net)
    desc += [int(has_bias), int(has_rdt)]
    flat = []
    for net in nets:
        for (W, b, r) in net:
            flat.append(np.asarray(W, np.float64).ravel())
            if has_bias:
                flat.append(np.asarray(b, np.float64).ravel())
            if r is not None:
                flat.append(np.asarray(r, np.float64).ravel())
    return np.asarray(desc, np.int32), np.concatenate(flat)


def pack_dp_weights(w):
    cfg = w.cfg
    ed, ep = _pack_nets(w.emb, [1] + list(cfg.emb_sizes))
    fd, fp = _pack_nets(w.fit, [cfg.emb_out * cfg.M2] + list(cfg.fit_sizes))
    return ed, ep, fd, fp


def dp_energy_grad(Ri, type_map, w, type_present=None):
    """Per-atom energy Ei [N] and dE = dEtot/dRi [N,S,4]."""
    cfg = w.cfg
    Ri = _c(Ri, np.float64)
    n = Ri.shape[0]
    if type_present is None:
        type_present = np.ones(cfg.ntypes, np.int32)
    ed, ep, fd, fp = pack_dp_weights(w)
    Ei = np.zeros(n)
    dE = np.zeros_like(Ri)
    rc = lib().oracle_dp_energy_grad(Ri, _c(type_map, np.int32), _c(type_present, np.int32), n,
                                     cfg.ntypes, cfg.m, cfg.M2, ed, ep, fd, fp, Ei, dE)
    if rc != 0:
        raise ValueError(f"oracle_dp_energy_grad failed ({rc})")
    return Ei, dE


def dp_force_virial(lst, dE, Ri_d, image_dr, nghost=0):
    lst = _c(lst, np.int32)
    n, S = lst.shape
    F = np.zeros((n + nghost, 3))
    V = np.zeros(9)
    lib().oracle_dp_force_virial(lst, _c(dE, np.float64), _c(Ri_d, np.float64),
                                 _c(image_dr, np.float64), n, nghost, S, F, V)
    return F, V


def dp_evaluate(frame, w, rc_nl=None, type_present=None):
    """Whole DP E/F/V of one frame, FP64: dict(E, Ei, F [N,3], V [9], list, dR, Ri, Ri_d, dE)."""
    cfg = w.cfg
    rc = cfg.rcut if rc_nl is None else rc_nl
    lst, dR = dp_find_neighbor(frame.lattice, frame.frac(), frame.types, cfg.ntypes, cfg.m, rc)
    idr = dp_image_dr(lst, dR)
    Ri, Ri_d = dp_env_mat(idr, frame.types, w.davg, w.dstd, cfg.rcut_smooth, cfg.rcut)
    Ei, dE = dp_energy_grad(Ri, frame.types, w, type_present)
    F, V = dp_force_virial(lst, dE, Ri_d, idr)
    return dict(E=float(Ei.sum()), Ei=Ei, F=F, V=V, list=lst, dR=dR, image_dr=idr, Ri=Ri,
                Ri_d=Ri_d, dE=dE)


# --------------------------------------------------------------------------- D7/D8 ops
def dp_force_op(nlist, dE, Ri_d, F_center, nghost=0):
    nlist = _c(nlist, np.int32)
    B, N, S = nlist.shape[0], nlist.shape[1], int(np.prod(nlist.shape[2:]))
    out = np.zeros((B, N + nghost, 3))
    lib().oracle_dp_force_op(nlist, _c(dE, np.float64), _c(Ri_d, np.float64),
                             _c(F_center, np.float64), B, N, S, nghost, out)
    return out


def dp_virial_op(nlist, dE, Rij, Ri_d, nghost=0):
    nlist = _c(nlist, np.int32)
    B, N, S = nlist.shape[0], nlist.shape[1], int(np.prod(nlist.shape[2:]))
    V = np.zeros((B, 9))
    AV = np.zeros((B, N + nghost, 9))
    lib().oracle_dp_virial_op(nlist, _c(dE, np.float64), _c(Rij, np.float64), _c(Ri_d, np.float64),
                              B, N, S, nghost, V, AV)
    return V, AV


def dp_force_grad_op(nlist, Ri_d, gF, nghost=0):
    nlist = _c(nlist, np.int32)
    B, N, S = nlist.shape[0], nlist.shape[1], int(np.prod(nlist.shape[2:]))
    g = np.zeros((B, N, S * 4))
    lib().oracle_dp_force_grad_op(nlist, _c(Ri_d, np.float64), _c(gF, np.float64), B, N, S, nghost, g)
    return g


def dp_virial_grad_op(nlist, Rij, Ri_d, gV):
    nlist = _c(nlist, np.int32)
    B, N, S = nlist.shape[0], nlist.shape[1], int(np.prod(nlist.shape[2:]))
    g = np.zeros((B, N, S * 4))
    lib().oracle_dp_virial_grad_op(nlist, _c(Rij, np.float64), _c(Ri_d, np.float64), _c(gV, np.float64),
                                   B, N, S, g)
    return g
