"""``torch.ops.CalcOps_cuda``: the reference operator library, re-implemented.

Registers the reference's op names and schemas (src/op/src/CalcOps_bind.cpp:6-17,
signatures src/op/include/CalcOps.h:35-359) on top of the C-ABI library: the
DP force/virial ops here, calculateCompress (matpl_amd.compress) and the NEP
ops (matpl_amd.nep_ops). The build's own entry points (cell-list neighbor
search, environment matrix, statistics, fused DP evaluation) are plain Python
functions in matpl_amd.neighbor / matpl_amd.dp_fused.

Conventions kept from the reference (src/op/src/CalcOps.cpp):
  * ops allocate and return their outputs as a list of tensors;
  * calculateForce returns [force (B, N+nghost, 3)] with rows < N set to the
    centre term passed in as F, and the neighbor term scattered on top;
  * calculateVirial returns [virial (B, 9), atom_virial (B, N+nghost, 9)];
  * gradients flow to dE only (CalcOps.cpp:25-39, 188-202).
Differences (the reference's behaviour is a bug or a host sync):
  * the scatter strides rows by N+nghost (the reference strides by N);
  * a wrong dtype/device raises instead of printing "data type error!";
  * nghost may be a CPU tensor (read without a device sync) or an int.
All work is enqueued on the current HIP stream.
"""
from __future__ import annotations

import torch

from . import _lib

_DTYPES = {torch.float32: "f32", torch.float64: "f64"}


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _ptr(t):
    return None if t is None else t.data_ptr()


def _check_gpu(*ts):
    for t in ts:
        if t is not None and not t.is_cuda:
            raise RuntimeError("CalcOps_cuda: tensors must be on the GPU")


def _aligned(t: torch.Tensor) -> torch.Tensor:
    t = t.contiguous()
    if t.data_ptr() % 16:
        t = t.clone()
    return t


def _sfx(t: torch.Tensor) -> str:
    if t.dtype not in _DTYPES:
        raise RuntimeError(f"CalcOps_cuda: unsupported dtype {t.dtype} (float32/float64)")
    return _DTYPES[t.dtype]


def _nghost(n) -> int:
    if isinstance(n, torch.Tensor):
        return int(n.item())
    return int(n)


# ----------------------------------------------------------------- raw launches
def dp_force(nlist, dE, Ri_d, F, nghost: int):
    _check_gpu(nlist, dE, Ri_d, F)
    B, N = nlist.shape[0], nlist.shape[1]
    S = nlist[0, 0].numel() if B * N > 0 else (nlist.numel() // max(B * N, 1))
    sfx = _sfx(dE)
    nlist = nlist.to(torch.int32).contiguous()
    dE = _aligned(dE)
    Ri_d = _aligned(Ri_d.to(dE.dtype))
    F = F.to(dE.dtype).contiguous()
    out = torch.empty((B, N + nghost, 3), dtype=dE.dtype, device=dE.device)
    _lib.call(f"matpl_dp_force_{sfx}", _ptr(nlist), _ptr(dE), _ptr(Ri_d), _ptr(F), B, N, S, nghost,
              _ptr(out), _stream())
    return out


def dp_virial(nlist, dE, Rij, Ri_d, nghost: int):
    _check_gpu(nlist, dE, Rij, Ri_d)
    B, N = nlist.shape[0], nlist.shape[1]
    S = nlist.numel() // max(B * N, 1)
    sfx = _sfx(dE)
    nlist = nlist.to(torch.int32).contiguous()
    dE = _aligned(dE)
    Ri_d = _aligned(Ri_d.to(dE.dtype))
    Rij = Rij.to(dE.dtype).contiguous()
    virial = torch.empty((B, 9), dtype=dE.dtype, device=dE.device)
    atom_virial = torch.empty((B, N + nghost, 9), dtype=dE.dtype, device=dE.device)
    ws_elems = _lib.query("matpl_dp_virial_workspace_elems", B, N, nghost)
    ws = torch.empty(max(ws_elems, 1), dtype=dE.dtype, device=dE.device)
    _lib.call(f"matpl_dp_virial_{sfx}", _ptr(nlist), _ptr(dE), _ptr(Rij), _ptr(Ri_d), B, N, S, nghost,
              _ptr(virial), _ptr(atom_virial), _ptr(ws), ws.numel(), _stream())
    return virial, atom_virial


def dp_force_grad(nlist, Ri_d, gF, nghost: int, dE_shape):
    B, N = nlist.shape[0], nlist.shape[1]
    S = nlist.numel() // max(B * N, 1)
    sfx = _sfx(Ri_d)
    g = torch.empty(dE_shape, dtype=Ri_d.dtype, device=Ri_d.device)
    # converted tensors are held in locals until the launch is enqueued: a
    # temporary freed earlier could be reused by the next conversion
    gF = gF.to(Ri_d.dtype).contiguous()
    nlist = nlist.to(torch.int32).contiguous()
    Ri_d = _aligned(Ri_d)
    _lib.call(f"matpl_dp_force_grad_{sfx}", _ptr(nlist), _ptr(Ri_d), _ptr(gF), B, N, S, nghost, _ptr(g),
              _stream())
    return g


def dp_virial_grad(nlist, Rij, Ri_d, gV, dE_shape):
    B, N = nlist.shape[0], nlist.shape[1]
    S = nlist.numel() // max(B * N, 1)
    sfx = _sfx(Ri_d)
    g = torch.empty(dE_shape, dtype=Ri_d.dtype, device=Ri_d.device)
    nlist = nlist.to(torch.int32).contiguous()
    Rij = Rij.to(Ri_d.dtype).contiguous()
    gV = gV.to(Ri_d.dtype).contiguous()
    Ri_d = _aligned(Ri_d)
    _lib.call(f"matpl_dp_virial_grad_{sfx}", _ptr(nlist), _ptr(Rij), _ptr(Ri_d), _ptr(gV), B, N, S, _ptr(g),
              _stream())
    return g


# ----------------------------------------------------------------- autograd wrappers
class CalculateForce(torch.autograd.Function):
    """CalcOps.cpp:41-62: forward scatters, backward returns grad wrt dE only."""

    @staticmethod
    def forward(ctx, list_neigh, dE, Ri_d, F, nghost_tensor):
        nghost = _nghost(nghost_tensor)
        ctx.save_for_backward(list_neigh, dE, Ri_d)
        ctx.nghost = nghost
        return dp_force(list_neigh, dE, Ri_d, F, nghost)

    @staticmethod
    def backward(ctx, gF):
        list_neigh, dE, Ri_d = ctx.saved_tensors
        g = dp_force_grad(list_neigh, Ri_d, gF, ctx.nghost, dE.shape)
        return None, g, None, None, None


class CalculateVirial(torch.autograd.Function):
    """CalcOps.cpp:204-226: [virial, atom_virial]; backward wrt dE from the virial grad."""

    @staticmethod
    def forward(ctx, list_neigh, dE, Rij, Ri_d, nghost_tensor):
        nghost = _nghost(nghost_tensor)
        ctx.save_for_backward(list_neigh, dE, Rij, Ri_d)
        v, av = dp_virial(list_neigh, dE, Rij, Ri_d, nghost)
        ctx.mark_non_differentiable(av)
        return v, av

    @staticmethod
    def backward(ctx, gV, gAV):
        list_neigh, dE, Rij, Ri_d = ctx.saved_tensors
        if gV is None:
            return None, None, None, None, None
        g = dp_virial_grad(list_neigh, Rij, Ri_d, gV, dE.shape)
        return None, g, None, None, None


def calculateForce(list_neigh, dE, Ri_d, F, nghost_tensor):
    return [CalculateForce.apply(list_neigh, dE, Ri_d, F, nghost_tensor)]


def calculateVirial(list_neigh, dE, Rij, Ri_d, nghost_tensor):
    v, av = CalculateVirial.apply(list_neigh, dE, Rij, Ri_d, nghost_tensor)
    return [v, av]


# ----------------------------------------------------------------- registration
_REGISTERED = False
_LIBS = []


def register() -> None:
    """Defines torch.ops.CalcOps_cuda.* (idempotent)."""
    global _REGISTERED
    if _REGISTERED:
        return
    lib = torch.library.Library("CalcOps_cuda", "DEF")
    lib.define("calculateForce(Tensor list_neigh, Tensor dE, Tensor Ri_d, Tensor F, Tensor nghost_tensor) -> Tensor[]")
    lib.define("calculateVirial(Tensor list_neigh, Tensor dE, Tensor Rij, Tensor Ri_d, Tensor nghost_tensor) -> Tensor[]")
    lib.impl("calculateForce", calculateForce, "CompositeImplicitAutograd")
    lib.impl("calculateVirial", calculateVirial, "CompositeImplicitAutograd")
    from . import compress, nep_ops
    lib.define("calculateCompress(Tensor f2, Tensor coefficient) -> Tensor[]")
    lib.impl("calculateCompress", compress.calculateCompress, "CompositeImplicitAutograd")
    # NEP schemas as bound in src/op/src/CalcOps_bind.cpp:10-16 (signatures src/op/include/CalcOps.h:168-318)
    lib.define("calculateNepFeat(Tensor coeff2, Tensor d12_radial, Tensor NL_radial, Tensor atom_map, Tensor feats, "
               "float rcut_radial, int feat_multi_nums) -> Tensor[]")
    lib.define("calculateNepMbFeat(Tensor coeff3, Tensor d12_angluar, Tensor NL_angluar, Tensor atom_map, "
               "Tensor feats, int feat_2b_num, int lmax_3, int lmax_4, int lmax_5, float rcut_angluar) -> Tensor[]")
    lib.define("calculateNepForce(Tensor list_neigh, Tensor dE, Tensor Ri_d, Tensor F) -> Tensor[]")
    lib.define("calculateNepVirial(Tensor list_neigh, Tensor dE, Tensor Rij, Tensor Ri_d, Tensor num_atom) "
               "-> Tensor[]")
    lib.define("calculate_maxneigh(Tensor num_atoms, Tensor box, Tensor box_orig, Tensor num_cell, Tensor position, "
               "float cutoff_2b, float cutoff_3b, int atom_type_num, Tensor atom_type_map, bool with_type=False) "
               "-> Tensor[]")
    lib.define("calculate_neighbor(Tensor num_atoms, Tensor atom_type_map, Tensor atom_types, Tensor box, "
               "Tensor box_orig, Tensor num_cell, Tensor position, float cutoff_2b, float cutoff_3b, "
               "int max_NN_radial, int max_NN_angular, bool with_rij=False) -> Tensor[]")
    lib.define("calculate_descriptor(Tensor weight_radial, Tensor weight_angular, Tensor Ri_radial, "
               "Tensor NL_radial, Tensor atom_type_map, float cutoff_radial, float cutoff_angular, "
               "int max_NN_radial, int lmax_3, int lmax_4, int lmax_5) -> Tensor[]")
    for name in ("calculateNepFeat", "calculateNepMbFeat", "calculateNepForce", "calculateNepVirial",
                 "calculate_maxneigh", "calculate_neighbor", "calculate_descriptor"):
        lib.impl(name, getattr(nep_ops, name), "CompositeImplicitAutograd")
    _LIBS.append(lib)
    _REGISTERED = True
