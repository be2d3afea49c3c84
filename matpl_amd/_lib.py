"""Loads libmatpl_hip.so (the C-ABI library) and checks every call.

The library is built in-tree by ``python -m matpl_amd.build``. There is no
CPU fallback: if the library is missing or a call fails, this raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmatpl_hip.so")
_LIB = None

# (name, restype, argtypes) for every exported symbol; must match include/matpl_hip.h
P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int
F64 = ctypes.c_double
SIGNATURES = {
    "matpl_last_error": (ctypes.c_char_p, []),
    "matpl_abi_version": (I32, []),
    # DP operator contract
    "matpl_dp_force_f32": (I32, [P, P, P, P, I64, I64, I64, I64, P, P]),
    "matpl_dp_force_f64": (I32, [P, P, P, P, I64, I64, I64, I64, P, P]),
    "matpl_dp_virial_workspace_elems": (I32, [I64, I64, I64]),
    "matpl_dp_virial_f32": (I32, [P, P, P, P, I64, I64, I64, I64, P, P, P, I64, P]),
    "matpl_dp_virial_f64": (I32, [P, P, P, P, I64, I64, I64, I64, P, P, P, I64, P]),
    "matpl_dp_virial_noalloc": (I32, [I32, P, P, P, P, I64, I64, I64, I64, P, P, P]),
    "matpl_dp_force_grad_f32": (I32, [P, P, P, I64, I64, I64, I64, P, P]),
    "matpl_dp_force_grad_f64": (I32, [P, P, P, I64, I64, I64, I64, P, P]),
    "matpl_dp_virial_grad_f32": (I32, [P, P, P, P, I64, I64, I64, P, P]),
    "matpl_dp_virial_grad_f64": (I32, [P, P, P, P, I64, I64, I64, P, P]),
    # DP neighbor search
    "matpl_dp_neighbor_is_small_box": (I32, [P, F64]),
    "matpl_dp_neighbor_cell_workspace_bytes": (I64, [P, F64, I64]),
    "matpl_dp_neighbor_brute": (I32, [P, P, P, P, P, P, I32, I32, F64, I64, I64, P, P, P, P, P, P, P]),
    "matpl_dp_neighbor_cell": (I32, [P, P, P, P, I32, I32, F64, I64, P, P, P, P, P, P, P, I64, P]),
    # DP environment matrix and statistics
    "matpl_dp_env_mat_f32": (I32, [P, P, P, P, I64, I64, I64, I32, F64, F64, P, P, P]),
    "matpl_dp_env_mat_f64": (I32, [P, P, P, P, I64, I64, I64, I32, F64, F64, P, P, P]),
    "matpl_dp_calc_stat_workspace_elems": (I64, [I32]),
    "matpl_dp_calc_stat": (I32, [P, P, P, I64, I64, I32, F64, F64, P, P, P, I64, P]),
    # fused DP descriptor + embedding net
    "matpl_dp_fused_forward": (I32, [P] * 12 + [I64, I32, I32, I32, I32, I32, F64, F64, F64, P, P, P]),
    "matpl_dp_fused_backward": (I32, [P] * 12 + [I64, I32, I32, I32, I32, I32, F64, F64, F64] + [P] * 8),
    "matpl_bsmm": (I32, [I32, I32, P, P, P, I64, I32, I32, I32, P]),
    "matpl_dp_train_tangent": (I32, [P] * 12 + [I64, I32, I32, I32, I32, I32, F64, F64, F64, P, P, P, P, P]),
    "matpl_dp_train_embgrad": (I32, [P] * 12 + [I64, I32, I32, I32, I32, I32, F64, F64, F64, P, P, P, P, I32, P, P]),
    "matpl_dp_train_grad_size": (I32, [I32]),
    "matpl_dp_fused_forward_compressed": (I32, [P] * 11 + [I32, I32, I32, F64, F64, P, I64, I32, I32, I32, I32, I32,
                                                          F64, F64, F64, P, P, P]),
    "matpl_dp_fused_backward_compressed": (I32, [P] * 11 + [I32, I32, I32, F64, F64, P, I64, I32, I32, I32, I32, I32,
                                                           F64, F64, F64] + [P] * 8),
    "matpl_dp_fused_gstore_words": (I64, [I64, I32, I32, I32, I32, I32]),
    "matpl_dp_fused_forward_store": (I32, [P] * 12 + [I64, I32, I32, I32, I32, I32, F64, F64, F64, P, P, P, P]),
    "matpl_dp_fused_backward_stored": (I32, [P] * 12 + [I64, I32, I32, I32, I32, I32, F64, F64, F64] + [P] * 9),
    "matpl_dp_fused_set_launch": (I32, [I32, I32, I32, I32]),
    "matpl_dp_fused_set_waves": (I32, [I32, I32]),
    "matpl_dp_fused_gather": (I32, [P, P, P, P, I32, I32, I64, P]),
    "matpl_dp_lmp_neighdata": (I32, [P, P, P, I32, P, P, P, P, I64, I64, I64, I32, I32, F64, P, P, P, P, P, P, P]),
    "matpl_dp_neighbor_reverse": (I32, [P, P, P, P, P, P, I32, I32, I64, P, P, P]),
    "matpl_dp_fitting": (I32, [P, P, I32, P, P, P, P, P, I64, I64, I64, I64, I32, I32, I32, P, P, P]),
    "matpl_dp_fitting_perm": (I32, [P, P, P, I32, P, P, P, P, P, I64, I64, I64, I64, I32, I32, I32, P, P, P]),
    "matpl_dp_fit_train": (I32, [P, P, P, P, P, I32, P, P, P, P, P, I64, I64, I64, I64, I32, I32, I32] + [P] * 9),
    "matpl_tall_xtz_work": (I64, [I64, I32, I32, I32]),
    "matpl_tall_xtz": (I32, [P, I64, I64, P, I64, I64, P, I32, I32, I32, P, P, I64, P]),
    "matpl_seg_rowsum_work": (I64, [I64, I32, I32]),
    "matpl_seg_rowsum": (I32, [I32, P, P, P, P, I32, P, P, I64, P]),
    # DP compress
    "matpl_dp_compress_f32": (I32, [P, P, I64, I32, I32, P, P]),
    "matpl_dp_compress_f64": (I32, [P, P, I64, I32, I32, P, P]),
    "matpl_dp_compress_grad_f32": (I32, [P, P, P, I64, I32, I32, P, P]),
    "matpl_dp_compress_grad_f64": (I32, [P, P, P, I64, I32, I32, P, P]),
    # NEP operators
    "matpl_nep_neighbor": (I32, [P] * 7 + [F64, F64, I64, I64, I64, I32, I32, I32] + [P] * 8),
    "matpl_nep_neighbor_ri32": (I32, [P] * 6 + [F64, F64, I64, I64, I64, I32] + [P] * 8),
    "matpl_nep_feat": (I32, [I32, I32, I32, P, P, P, P, I64] + [I32] * 9 + [F64] + [P] * 8),
    "matpl_nep_force": (I32, [I32, P, P, P, P, I64, I32, P]),
    "matpl_nep_virial": (I32, [I32, P, P, P, P, P, I64, I64, I32, P, P, P]),
    "matpl_nep_centre_force": (I32, [I32, P, P, I64, I32, P, P]),
    "matpl_nep_frame_virial": (I32, [I32, P, P, P, P, P, I64, I64, I32, P, P, P]),
    "matpl_nep_force_grad": (I32, [I32, P, P, P, P, I64, I32, P]),
    "matpl_nep_virial_grad": (I32, [I32, P, P, P, P, P, P, I64, I32, P]),
    "matpl_nep_force_virial_ri_work": (I64, [I32, I64, I64]),
    "matpl_nep_force_virial_ri": (I32, [I32, P, P, P, P, P, I64, I64, I32, I64, P, P, P, I64, P]),
    "matpl_nep_force_virial_ri_grad": (I32, [I32, P, P, P, P, P, I64, I32, P, P]),
    "matpl_nep_fit_tile_atoms": (I32, [I32, I32, I32]),
    "matpl_nep_fit_part_width": (I32, [I32, I32]),
    "matpl_nep_fit_forward": (I32, [I32, P, P, P, P, I32, I32, P, P, P, P, I32, I32, P, P, P]),
    "matpl_nep_fit_backward": (I32, [I32, P, P, P, P, I32, I32, P, I32, P, P, P, P, I32, I32, P, P, P, P, P, P, P, P,
                                     P]),
    # MD helpers
    "matpl_md_kick": (I32, [P, P, P, P, F64, I64, I32, P]),
    "matpl_md_affine": (I32, [P, P, P, I64, P]),
    "matpl_tall_mm": (I32, [P, I64, P, I64, I64, P, I64, I32, I32, I32, P]),
}


class MatplError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise MatplError(f"{LIB_PATH} is missing; build it with `python -m matpl_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    return list(SIGNATURES)


def call(name: str, *args) -> int:
    """Calls an entry point returning a status code; raises on failure."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().matpl_last_error().decode(errors="replace")
        raise MatplError(f"{name} failed ({rc}): {msg}")
    return rc


def query(name: str, *args):
    """Calls an entry point that returns a value (sizes, flags)."""
    return getattr(lib(), name)(*args)
