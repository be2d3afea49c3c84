"""NEP operators of ``torch.ops.CalcOps_cuda`` on the HIP library.

Autograd structure follows the reference (src/op/src/CalcOps.cpp:385-1008):
feature op -> its gradient op (grad wrt the coefficients and the [N, S, 4]
[r, x, y, z] inputs) -> the gradient op's backward (second order wrt the
upstream feature gradient and the coefficients), so that a force-loss
backward reaches the fitting-net and descriptor parameters.

Differences from the reference, all where the reference is a bug:
  * float32 and float64 (the reference casts to double without a check);
  * the angular gradient uses the true rcut_angular (the reference passes it
    as an int, calculate_nepfeat.h:139);
  * the virial backward uses each atom's own frame gradient (the reference
    applies frame 0's, calculateNepVirial.cu:110-157);
  * the second-order pass accumulates in the op's dtype (the reference has a
    float accumulator in an FP64 kernel, calculateNepMbFeat_secondgradout.cu:20);
  * outputs are new tensors; the feature ops do not write into ``feats`` and
    calculateNepForce returns a new tensor instead of mutating ``F``.
As in the reference, the second-order pass ignores the gradient flowing into
the coefficient-gradient output (zero in force-loss training).
"""
from __future__ import annotations

from contextlib import contextmanager
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .calcops import _ptr, _stream

_F64 = {torch.float64: 1, torch.float32: 0}


def _f64(t):
    if t.dtype not in _F64:
        raise RuntimeError(f"CalcOps_cuda NEP ops: unsupported dtype {t.dtype}")
    return _F64[t.dtype]


def _nangular(na, l3, l4, l5):
    return na * l3 + (na if l4 > 0 else 0) + (na if l5 > 0 else 0)


def _feat_call(mode, angular, c, Ri, NL, atom_map, rc, l=(0, 0, 0), feats=None, s_save=None, gfeat=None, gg=None,
               gRi=None, gc=None, ggfeat=None):
    N, S = NL.shape
    T, _, nm, nb = c.shape
    ld = _nangular(nm, *l) if angular else nm
    _lib.call("matpl_nep_feat", _f64(Ri), int(angular), mode, _ptr(c), _ptr(Ri), _ptr(NL), _ptr(atom_map), N, S, T,
              nm, nb, ld, 0, int(l[0]), int(l[1]), int(l[2]), float(rc), _ptr(feats), _ptr(s_save), _ptr(gfeat),
              _ptr(gg), _ptr(gRi), _ptr(gc), _ptr(ggfeat), _stream())


def _expect(t, numel, what):
    if t.numel() != numel:
        raise ValueError(f"CalcOps_cuda NEP ops: {what} has {t.numel()} elements, expected {numel}")


@dataclass
class StaticFrames:
    num_atom: torch.Tensor
    start: torch.Tensor      # [F + 1] int64, first atom of each frame
    af: torch.Tensor         # [N] int32, frame of each atom
    status: torch.Tensor     # [1] int32, neighbor-overflow flag of the ops called inside (sticky)
    max_frame: int = 0       # atoms in the largest frame


_STATIC = None


@contextmanager
def static_frames(num_atom: torch.Tensor):
    """Inside this context the NEP ops make no host synchronisation, so that a sequence of them can be
    captured in a HIP graph: the frame index of ``num_atom`` (this tensor object) is computed once on
    entry, input checks that read device values are skipped, and calculate_neighbor records overflow in
    the context's ``status`` (1 + atom, sticky) instead of raising. The caller runs the same inputs
    through the checked ops first and reads ``status`` after the work."""
    global _STATIC
    start, af = _frame_index(num_atom, num_atom.device)
    ctx = StaticFrames(num_atom, start, af, torch.zeros(1, dtype=torch.int32, device=num_atom.device),
                       max(_int_max(num_atom), 1))
    prev = _STATIC
    _STATIC = ctx
    try:
        yield ctx
    finally:
        _STATIC = prev


# Host-side input checks read device values (a synchronization each). Their results are kept on the
# checked tensor, keyed by its version counter (bumped by every in-place change), so that a training
# loop passing the same num_atom / atom_type_map every step, and the lists that calculate_neighbor
# builds (valid by construction), pay for each check once.
def _cached(t, key, fn):
    try:
        ver = t._version
    except RuntimeError:           # inference-mode tensors keep no version counter: no caching
        return fn()
    hit = getattr(t, "_matpl_checks", None)
    if hit is not None and hit[0] == ver and key in hit[1]:
        return hit[1][key]
    if hit is None or hit[0] != ver:
        hit = (ver, {})
        t._matpl_checks = hit
    val = fn()
    hit[1][key] = val
    return val


def _int_max(t):
    """The largest entry (for the lists calculate_neighbor builds, its bound N - 1)."""
    return _cached(t, "max", lambda: int(t.max().item()) if t.numel() else -1)


def _int_sum(t):
    return _cached(t, "sum", lambda: int(t.sum().item()))


def _check_list(NL, natoms):
    """Neighbor indices must address atoms of the packed batch (the kernels index with them)."""
    if _STATIC is None and NL.numel() and _int_max(NL) >= natoms:
        raise ValueError(f"CalcOps_cuda NEP ops: neighbor index {_int_max(NL)} >= {natoms} atoms")


def _prep(c, Ri, NL, atom_map):
    Ri = Ri.contiguous()
    NL = NL.to(torch.int64).contiguous()
    atom_map = atom_map.to(torch.int64).contiguous()
    if c.dim() != 4 or c.shape[0] != c.shape[1]:
        raise ValueError(f"CalcOps_cuda NEP ops: coefficients must be [T, T, n, k], got {tuple(c.shape)}")
    if NL.dim() != 2 or atom_map.numel() != NL.shape[0]:
        raise ValueError("CalcOps_cuda NEP ops: NL must be [N, S] with one atom_map entry per atom")
    _expect(Ri, NL.numel() * 4, "Ri [N, S, 4]")
    _check_list(NL, NL.shape[0])
    if _STATIC is None and atom_map.numel() and _int_max(atom_map) >= c.shape[0]:
        raise ValueError("CalcOps_cuda NEP ops: atom type beyond the coefficient table")
    return (c.to(Ri.dtype).contiguous(), Ri, NL, atom_map)


# ------------------------------------------------------------------ radial
class NepFeatRadial(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c2, Ri, NL, atom_map, rc):
        c2, Ri, NL, atom_map = _prep(c2, Ri, NL, atom_map)
        feats = torch.empty((NL.shape[0], c2.shape[2]), dtype=Ri.dtype, device=Ri.device)
        _feat_call(0, False, c2, Ri, NL, atom_map, rc, feats=feats)
        ctx.save_for_backward(c2, Ri, NL, atom_map)
        ctx.rc = rc
        return feats

    @staticmethod
    def backward(ctx, g):
        c2, Ri, NL, atom_map = ctx.saved_tensors
        gc2, gRi = NepFeatRadialGrad.apply(g.contiguous(), c2, Ri, NL, atom_map, ctx.rc)
        return gc2, gRi, None, None, None


class NepFeatRadialGrad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, g, c2, Ri, NL, atom_map, rc):
        gRi = torch.empty(Ri.shape[:2] + (4,), dtype=Ri.dtype, device=Ri.device)
        gc2 = torch.zeros_like(c2)
        _feat_call(1, False, c2, Ri, NL, atom_map, rc, gfeat=g, gRi=gRi, gc=gc2)
        ctx.save_for_backward(g, c2, Ri, NL, atom_map)
        ctx.rc = rc
        return gc2, gRi

    @staticmethod
    def backward(ctx, ggc2, ggRi):
        g, c2, Ri, NL, atom_map = ctx.saved_tensors
        if ggRi is None:
            return None, None, None, None, None, None
        ggfeat = torch.empty_like(g)
        gc2 = torch.zeros_like(c2)
        _feat_call(2, False, c2, Ri, NL, atom_map, ctx.rc, gfeat=g, gg=ggRi.contiguous(), gc=gc2, ggfeat=ggfeat)
        return ggfeat, gc2, None, None, None, None


# ------------------------------------------------------------------ angular
class NepFeatAngular(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c3, Ri, NL, atom_map, rc, l3, l4, l5):
        c3, Ri, NL, atom_map = _prep(c3, Ri, NL, atom_map)
        nm = c3.shape[2]
        feats = torch.empty((NL.shape[0], _nangular(nm, l3, l4, l5)), dtype=Ri.dtype, device=Ri.device)
        s_save = torch.empty((NL.shape[0], nm * 24), dtype=Ri.dtype, device=Ri.device)
        _feat_call(0, True, c3, Ri, NL, atom_map, rc, (l3, l4, l5), feats=feats, s_save=s_save)
        ctx.save_for_backward(c3, Ri, NL, atom_map, s_save)
        ctx.cfg = (rc, l3, l4, l5)
        return feats

    @staticmethod
    def backward(ctx, g):
        c3, Ri, NL, atom_map, s_save = ctx.saved_tensors
        gc3, gRi = NepFeatAngularGrad.apply(g.contiguous(), c3, Ri, NL, atom_map, s_save, *ctx.cfg)
        return gc3, gRi, None, None, None, None, None, None


class NepFeatAngularGrad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, g, c3, Ri, NL, atom_map, s_save, rc, l3, l4, l5):
        gRi = torch.empty(Ri.shape[:2] + (4,), dtype=Ri.dtype, device=Ri.device)
        gc3 = torch.zeros_like(c3)
        _feat_call(1, True, c3, Ri, NL, atom_map, rc, (l3, l4, l5), s_save=s_save, gfeat=g, gRi=gRi, gc=gc3)
        ctx.save_for_backward(g, c3, Ri, NL, atom_map, s_save)
        ctx.cfg = (rc, l3, l4, l5)
        return gc3, gRi

    @staticmethod
    def backward(ctx, ggc3, ggRi):
        g, c3, Ri, NL, atom_map, s_save = ctx.saved_tensors
        if ggRi is None:
            return (None,) * 10
        rc, l3, l4, l5 = ctx.cfg
        ggfeat = torch.empty_like(g)
        gc3 = torch.zeros_like(c3)
        _feat_call(2, True, c3, Ri, NL, atom_map, rc, (l3, l4, l5), s_save=s_save, gfeat=g, gg=ggRi.contiguous(),
                   gc=gc3, ggfeat=ggfeat)
        return ggfeat, gc3, None, None, None, None, None, None, None, None


# ------------------------------------------------------------------ force / virial
class NepForce(torch.autograd.Function):
    """calculateNepForce: F + neighbor scatter (CalcOps.cpp:75-158); grad wrt dE only."""

    @staticmethod
    def forward(ctx, NL, dE, Ri_d, F):
        NL = NL.to(torch.int64).contiguous()
        dE = dE.contiguous()
        Ri_d = Ri_d.to(dE.dtype).contiguous()
        out = F.to(dE.dtype).contiguous().clone()
        N, S = NL.shape
        _expect(dE, N * S * 4, "dE [N, S, 4]")
        _expect(Ri_d, N * S * 12, "Ri_d [N, S, 4, 3]")
        _expect(out, N * 3, "F [N, 3]")
        _check_list(NL, N)
        _lib.call("matpl_nep_force", _f64(dE), _ptr(NL), _ptr(dE), _ptr(Ri_d), _ptr(out), N, S, _stream())
        ctx.save_for_backward(NL, Ri_d)
        ctx.dE_shape = dE.shape
        return out

    @staticmethod
    def backward(ctx, gF):
        NL, Ri_d = ctx.saved_tensors
        N, S = NL.shape
        g = torch.empty(ctx.dE_shape, dtype=Ri_d.dtype, device=Ri_d.device)
        gF = gF.to(Ri_d.dtype).contiguous()
        _lib.call("matpl_nep_force_grad", _f64(Ri_d), _ptr(NL), _ptr(Ri_d), _ptr(gF), _ptr(g), N, S, _stream())
        return None, g, None, None


def _frame_index(num_atom, device):
    if _STATIC is not None and num_atom is _STATIC.num_atom:
        return _STATIC.start, _STATIC.af

    def build():
        na = num_atom.reshape(-1).to(torch.int64)
        start = torch.zeros(na.numel() + 1, dtype=torch.int64, device=device)
        start[1:] = torch.cumsum(na.to(device), 0)
        af = torch.repeat_interleave(torch.arange(na.numel(), device=device, dtype=torch.int32), na.to(device))
        return start, af
    return _cached(num_atom, ("frames", str(device)), build)


class NepVirial(torch.autograd.Function):
    """calculateNepVirial: [virial per frame, atom_virial] (CalcOps.cpp:239-328)."""

    @staticmethod
    def forward(ctx, NL, dE, Rij, Ri_d, num_atom):
        NL = NL.to(torch.int64).contiguous()
        dE = dE.contiguous()
        Rij = Rij.to(dE.dtype).contiguous()
        Ri_d = Ri_d.to(dE.dtype).contiguous()
        N, S = NL.shape
        _expect(dE, N * S * 4, "dE [N, S, 4]")
        _expect(Ri_d, N * S * 12, "Ri_d [N, S, 4, 3]")
        _expect(Rij, N * S * 3, "Rij [N, S, 3]")
        _check_list(NL, N)
        if _STATIC is None and _int_sum(num_atom) != N:
            raise ValueError(f"calculateNepVirial: num_atom sums to {_int_sum(num_atom)}, list has {N} atoms")
        start, af = _frame_index(num_atom, dE.device)
        nf = start.numel() - 1
        virial = torch.empty((nf, 9), dtype=dE.dtype, device=dE.device)
        av = torch.empty((N, 9), dtype=dE.dtype, device=dE.device)
        _lib.call("matpl_nep_virial", _f64(dE), _ptr(NL), _ptr(dE), _ptr(Rij), _ptr(Ri_d), _ptr(start), nf, N, S,
                  _ptr(virial), _ptr(av), _stream())
        ctx.save_for_backward(NL, Rij, Ri_d, af)
        ctx.dE_shape = dE.shape
        ctx.mark_non_differentiable(av)
        return virial, av

    @staticmethod
    def backward(ctx, gV, gAV):
        NL, Rij, Ri_d, af = ctx.saved_tensors
        if gV is None:
            return None, None, None, None, None
        N, S = NL.shape
        g = torch.empty(ctx.dE_shape, dtype=Ri_d.dtype, device=Ri_d.device)
        gV = gV.to(Ri_d.dtype).contiguous()
        _lib.call("matpl_nep_virial_grad", _f64(Ri_d), _ptr(NL), _ptr(Rij), _ptr(Ri_d), _ptr(gV), _ptr(af), _ptr(g),
                  N, S, _stream())
        return None, g, None, None, None


class NepFrameVirial(torch.autograd.Function):
    """The per-frame virial of calculateNepVirial without the per-atom rows (what the NEP model
    keeps, src/model/nep_net.py:613-622); same backward as NepVirial."""

    @staticmethod
    def forward(ctx, NL, dE, Rij, Ri_d, num_atom):
        NL = NL.to(torch.int64).contiguous()
        dE = dE.contiguous()
        Rij = Rij.to(dE.dtype).contiguous()
        Ri_d = Ri_d.to(dE.dtype).contiguous()
        N, S = NL.shape
        _expect(dE, N * S * 4, "dE [N, S, 4]")
        _expect(Ri_d, N * S * 12, "Ri_d [N, S, 4, 3]")
        _expect(Rij, N * S * 3, "Rij [N, S, 3]")
        start, af = _frame_index(num_atom, dE.device)
        if af.numel() != N:
            raise ValueError(f"nep frame virial: num_atom sums to {af.numel()}, list has {N} atoms")
        nf = start.numel() - 1
        virial = torch.empty((nf, 9), dtype=dE.dtype, device=dE.device)
        rows = torch.empty((N, 9), dtype=dE.dtype, device=dE.device)
        _lib.call("matpl_nep_frame_virial", _f64(dE), _ptr(NL), _ptr(dE), _ptr(Rij), _ptr(Ri_d), _ptr(start), nf, N,
                  S, _ptr(rows), _ptr(virial), _stream())
        ctx.save_for_backward(NL, Rij, Ri_d, af)
        ctx.dE_shape = dE.shape
        return virial

    @staticmethod
    def backward(ctx, gV):
        NL, Rij, Ri_d, af = ctx.saved_tensors
        N, S = NL.shape
        g = torch.empty(ctx.dE_shape, dtype=Ri_d.dtype, device=Ri_d.device)
        gV = gV.to(Ri_d.dtype).contiguous()
        _lib.call("matpl_nep_virial_grad", _f64(Ri_d), _ptr(NL), _ptr(Rij), _ptr(Ri_d), _ptr(gV), _ptr(af), _ptr(g),
                  N, S, _stream())
        return None, g, None, None, None


def nep_frame_virial(NL, dE, Rij, Ri_d, num_atom):
    return NepFrameVirial.apply(NL, dE, Rij, Ri_d, num_atom)


class NepForceVirialRi(torch.autograd.Function):
    """The NEP model's force and frame virial of one list from dE and Ri = [r, x, y, z], with the position
    derivative formed in the kernel: the sequence ri_d (src/model/nep_net.py:440-468), centre term
    (nep_net.py:608), calculateNepForce (CalcOps.cpp:75-158) and the frame sum of calculateNepVirial
    (CalcOps.cpp:239-328, nep_net.py:613-622) in one pass. Returns F [N, 3] (whole force, the
    reference's sign before the model's negation) and V [frames, 9]; differentiable in dE only (Ri is
    taken as a constant, as the model detaches it for Ri_d)."""

    @staticmethod
    def forward(ctx, NL, dE, Ri, num_atom):
        NL = NL.to(torch.int64).contiguous()
        dE = dE.contiguous()
        Ri = Ri.detach().to(dE.dtype).contiguous()
        N, S = NL.shape
        _expect(dE, N * S * 4, "dE [N, S, 4]")
        _expect(Ri, N * S * 4, "Ri [N, S, 4]")
        _check_list(NL, N)
        if _STATIC is not None and num_atom is _STATIC.num_atom:
            maxf = _STATIC.max_frame
        else:
            if _int_sum(num_atom) != N:
                raise ValueError(f"nep force/virial: num_atom sums to {_int_sum(num_atom)}, list has {N} atoms")
            maxf = max(_int_max(num_atom), 1)
        start, af = _frame_index(num_atom, dE.device)
        nf = start.numel() - 1
        F = torch.empty((N, 3), dtype=dE.dtype, device=dE.device)
        V = torch.empty((nf, 9), dtype=dE.dtype, device=dE.device)
        work = torch.empty(max(_lib.query("matpl_nep_force_virial_ri_work", _f64(dE), N, maxf), 1), dtype=dE.dtype,
                           device=dE.device)
        _lib.call("matpl_nep_force_virial_ri", _f64(dE), _ptr(NL), _ptr(dE), _ptr(Ri), _ptr(start), _ptr(af), nf, N,
                  S, maxf, _ptr(F), _ptr(V), _ptr(work), work.numel(), _stream())
        ctx.save_for_backward(NL, Ri, af)
        ctx.dE_shape = dE.shape
        return F, V

    @staticmethod
    def backward(ctx, gF, gV):
        NL, Ri, af = ctx.saved_tensors
        N, S = NL.shape
        if gF is None and gV is None:
            return None, None, None, None
        g = torch.empty(ctx.dE_shape, dtype=Ri.dtype, device=Ri.device)
        gF = None if gF is None else gF.to(Ri.dtype).contiguous()
        gV = None if gV is None else gV.to(Ri.dtype).contiguous()
        _lib.call("matpl_nep_force_virial_ri_grad", _f64(Ri), _ptr(NL), _ptr(Ri), _ptr(gF), _ptr(gV), _ptr(af), N, S,
                  _ptr(g), _stream())
        return None, g, None, None


def nep_force_virial(NL, dE, Ri, num_atom):
    """(F, V) of one neighbor list from dE and Ri; see NepForceVirialRi."""
    return NepForceVirialRi.apply(NL, dE, Ri, num_atom)


def nep_centre_force(dE, Ri_d):
    """-sum_{k,w} dE[i,k,w] Ri_d[i,k,w,:] -> [N, 3] (the centre term src/model/nep_net.py:608 forms with
    torch.matmul). Not differentiable: calculateNepForce's backward carries the centre term itself and
    returns no gradient for F."""
    N = dE.shape[0]
    S = dE.numel() // max(4 * N, 1)
    _expect(Ri_d, N * S * 12, "Ri_d [N, S, 4, 3]")
    dE = dE.detach().contiguous()
    Ri_d = Ri_d.detach().to(dE.dtype).contiguous()
    F = torch.empty((N, 3), dtype=dE.dtype, device=dE.device)
    _lib.call("matpl_nep_centre_force", _f64(dE), _ptr(dE), _ptr(Ri_d), N, S, _ptr(F), _stream())
    return F


# ------------------------------------------------------------------ neighbor search
def _nep_geo(num_atoms, box, box_orig, num_cell, position):
    dev = position.device
    nf = num_atoms.numel()
    if (_STATIC is None and _int_sum(num_atoms) != position.shape[0]) or \
            position.numel() != 3 * position.shape[0]:
        raise ValueError("NEP neighbor search: position must be [sum(num_atoms), 3]")
    if box.numel() != 18 * nf or box_orig.numel() != 9 * nf or num_cell.numel() != 3 * nf:
        raise ValueError("NEP neighbor search: box [F*18], box_orig [F*9], num_cell [F*3] expected")
    start, af = _frame_index(num_atoms, dev)
    return (start, af, box.to(torch.float64).contiguous(), box_orig.to(torch.float64).contiguous(),
            num_cell.to(torch.int64).contiguous(), position.to(torch.float64).contiguous())


def calculate_maxneigh(num_atoms, box, box_orig, num_cell, position, cutoff_2b, cutoff_3b, atom_type_num,
                       atom_type_map, with_type=False):
    start, af, box, bo, nc, pos = _nep_geo(num_atoms, box, box_orig, num_cell, position)
    N = pos.shape[0]
    if with_type and (atom_type_map.numel() != N or
                      (_STATIC is None and N and _int_max(atom_type_map) >= atom_type_num)):
        raise ValueError("calculate_maxneigh: atom_type_map must give each atom a type < atom_type_num")
    shape = (N, atom_type_num) if with_type else (N,)
    NN_r = torch.empty(shape, dtype=torch.int64, device=pos.device)
    NN_a = torch.empty(shape, dtype=torch.int64, device=pos.device)
    tm = atom_type_map.to(torch.int64).contiguous()
    _lib.call("matpl_nep_neighbor", _ptr(box), _ptr(bo), _ptr(nc), _ptr(start), _ptr(af), _ptr(tm), _ptr(pos),
              float(cutoff_2b), float(cutoff_3b), 0, 0, N, 2 if with_type else 1, int(atom_type_num), 1, _ptr(NN_r),
              None, _ptr(NN_a), None, None, None, None, _stream())
    return [NN_r.to(num_atoms.dtype), NN_a.to(num_atoms.dtype)]


def calculate_neighbor(num_atoms, atom_type_map, atom_types, box, box_orig, num_cell, position, cutoff_2b, cutoff_3b,
                       max_NN_radial, max_NN_angular, with_rij=False, ri_dtype=None):
    """[NN_radial, NN_angular, NL_radial, NL_angular, Ri_radial, Ri_angular] (CalcOps.cpp:1089-1172).
    ri_dtype=torch.float32 (not in the reference's signature): the Ri rows in float, rounded once from
    the FP64 values in the search kernel (equal to casting the default rows)."""
    start, af, box, bo, nc, pos = _nep_geo(num_atoms, box, box_orig, num_cell, position)
    N = pos.shape[0]
    dev = pos.device
    w = 4 if with_rij else 3
    ri32 = ri_dtype == torch.float32
    if ri_dtype is not None and ri_dtype not in (torch.float32, torch.float64):
        raise ValueError("calculate_neighbor: ri_dtype must be float32 or float64")
    NN_r = torch.empty(N, dtype=torch.int64, device=dev)
    NN_a = torch.empty(N, dtype=torch.int64, device=dev)
    NL_r = torch.empty((N, max_NN_radial), dtype=torch.int64, device=dev)
    NL_a = torch.empty((N, max_NN_angular), dtype=torch.int64, device=dev)
    rdt = torch.float32 if ri32 else torch.float64
    Ri_r = torch.empty((N, max_NN_radial, w), dtype=rdt, device=dev)
    Ri_a = torch.empty((N, max_NN_angular, w), dtype=rdt, device=dev)
    status = _STATIC.status if _STATIC is not None else torch.zeros(1, dtype=torch.int32, device=dev)
    if ri32:
        _lib.call("matpl_nep_neighbor_ri32", _ptr(box), _ptr(bo), _ptr(nc), _ptr(start), _ptr(af), _ptr(pos),
                  float(cutoff_2b), float(cutoff_3b), int(max_NN_radial), int(max_NN_angular), N, int(with_rij),
                  _ptr(NN_r), _ptr(NL_r), _ptr(NN_a), _ptr(NL_a), _ptr(Ri_r), _ptr(Ri_a), _ptr(status), _stream())
    else:
        _lib.call("matpl_nep_neighbor", _ptr(box), _ptr(bo), _ptr(nc), _ptr(start), _ptr(af), None, _ptr(pos),
                  float(cutoff_2b), float(cutoff_3b), int(max_NN_radial), int(max_NN_angular), N, 0, 0,
                  int(with_rij), _ptr(NN_r), _ptr(NL_r), _ptr(NN_a), _ptr(NL_a), _ptr(Ri_r), _ptr(Ri_a), _ptr(status),
                  _stream())
    s = int(status.item()) if _STATIC is None else 0
    if s:
        raise OverflowError(f"calculate_neighbor: more than max_NN neighbors (atom {s - 1})")
    it = num_atoms.dtype
    dt = ri_dtype if ri_dtype is not None else position.dtype
    out = [NN_r.to(it), NN_a.to(it), NL_r.to(it), NL_a.to(it), Ri_r.to(dt), Ri_a.to(dt)]
    # the lists address atoms of this batch only (-1 marks an empty slot): record it for the ops' checks
    for NL in out[2:4]:
        _cached(NL, "max", lambda: N - 1)
    return out


def calculate_descriptor(weight_radial, weight_angular, Ri_radial, NL_radial, atom_type_map, cutoff_radial,
                         cutoff_angular, max_NN_radial, lmax_3, lmax_4, lmax_5):
    """Features from the radial list ([x,y,z] or [r,x,y,z]), angular part within cutoff_angular
    (CalcOps.cpp:1174-1230, nep3_small_box.cuh:368-463)."""
    Ri = Ri_radial
    if Ri.shape[-1] == 3:
        x, y, z = Ri.unbind(-1)
        Ri = torch.stack([torch.sqrt(x * x + y * y + z * z), x, y, z], -1)
    Ri = Ri.contiguous()
    NL = NL_radial.to(torch.int64).contiguous()
    am = atom_type_map.to(torch.int64)
    # the angular sums run over the radial list: neighbors beyond cutoff_angular have f_k = 0 and add nothing
    fr = NepFeatRadial.apply(weight_radial, Ri, NL, am, float(cutoff_radial))
    fa = NepFeatAngular.apply(weight_angular, Ri, NL, am, float(cutoff_angular), int(lmax_3), int(lmax_4),
                              int(lmax_5))
    return [torch.cat([fr, fa], -1)]


# ------------------------------------------------------------------ reference-named entry points
def calculateNepFeat(coeff2, d12_radial, NL_radial, atom_map, feats, rcut_radial, multi_feat_num):
    return [NepFeatRadial.apply(coeff2, d12_radial, NL_radial, atom_map, float(rcut_radial))]


def calculateNepMbFeat(coeff3, d12, NL, atom_map, feats, feat_2b_num, lmax_3, lmax_4, lmax_5, rcut_angular):
    return [NepFeatAngular.apply(coeff3, d12, NL, atom_map, float(rcut_angular), int(lmax_3), int(lmax_4),
                                 int(lmax_5))]


def calculateNepForce(list_neigh, dE, Ri_d, F):
    return [NepForce.apply(list_neigh, dE, Ri_d, F)]


def calculateNepVirial(list_neigh, dE, Rij, Ri_d, num_atom):
    v, av = NepVirial.apply(list_neigh, dE, Rij, Ri_d, num_atom)
    return [v, av]
