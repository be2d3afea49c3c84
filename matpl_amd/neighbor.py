"""DP neighbor search on the GPU (the find_neighbore contract).

``find_neighbore`` mirrors src/pre_data/dpuni_data_loader.py:595-650 for one
frame: it returns list_neigh [N, T*m] (1-based, 0 = empty) and dR [N, T*m, 3]
(float64, r_j - r_i), identical to NeighConst.f90:14-149. Large boxes use the
cell-list kernel, boxes with a lattice vector shorter than 2*rc use the
image-enumerating brute-force kernel (the reference's own small-box rule).
``find_neighbore_batch`` runs many frames in one launch.

An overflowing bucket (more than m neighbors of one type) raises
OverflowError, as the reference stops with "maxNeighborNum too small".

``generate_neighdata`` is the LAMMPS-side packing (PairMATPL::generate_neighdata)
from a full neighbor list over local + ghost atoms, on the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from . import _lib
from .calcops import _ptr, _stream


@dataclass
class NeighborList:
    list: torch.Tensor          # [Ntot, T*m] int32, 1-based within frame, 0 = empty
    dR: Optional[torch.Tensor]  # [Ntot, T*m, 3] float64
    counts: torch.Tensor        # [Ntot, T] int32
    shift: Optional[torch.Tensor]  # [Ntot, T*m] int32 packed image shift
    frac_wrapped: torch.Tensor  # [Ntot, 3] float64
    status: torch.Tensor        # [1] int32: 0, 1 + first overflowing atom, or STATUS_NONFINITE

    def check(self):
        s = int(self.status.item())
        if s == STATUS_NONFINITE:
            raise ValueError("neighbor search: a position is not finite (NaN or inf)")
        if s != 0:
            raise OverflowError(f"maxNeighborNum too small (atom {s - 1})")
        return self


STATUS_NONFINITE = -4           # csrc/neighbor.hip: a coordinate is NaN or infinite


# small device-side constants (cutoffs per type, one-frame index arrays) are made once and
# reused, so that repeated calls - including under stream capture - do no host-to-device copies
_CONST = {}


def _const(key, make):
    t = _CONST.get(key)
    if t is None:
        t = _CONST[key] = make()
    return t


def _rc_type_tensor(rc_type, ntypes, rc_m, device):
    if rc_type is None:
        rc_type = np.full(ntypes, rc_m)
    a = np.asarray(rc_type, np.float64)
    return _const(("rct", str(device), a.tobytes()), lambda: torch.as_tensor(a, device=device))


def find_neighbore_gpu(lattice, frac, type_map, ntypes: int, m: int, rc_m: float, rc_type=None,
                       want_dR=True, want_shift=False, device="cuda", force_brute=False,
                       check=True, out: Optional[NeighborList] = None) -> NeighborList:
    """One frame. lattice [3,3] (host array; rows are cell vectors); frac [N,3].

    ``out``: preallocated result buffers (same shapes) to write into.
    """
    lat = np.ascontiguousarray(np.asarray(lattice, np.float64).reshape(3, 3))
    frac_t = torch.as_tensor(frac, dtype=torch.float64, device=device).contiguous()
    tm = torch.as_tensor(type_map, dtype=torch.int32, device=device).contiguous()
    n = frac_t.shape[0]
    S = ntypes * m
    rct = _rc_type_tensor(rc_type, ntypes, rc_m, device)
    if out is None:
        out = NeighborList(
            list=torch.empty((n, S), dtype=torch.int32, device=device),
            dR=torch.empty((n, S, 3), dtype=torch.float64, device=device) if want_dR else None,
            counts=torch.empty((n, ntypes), dtype=torch.int32, device=device),
            shift=torch.empty((n, S), dtype=torch.int32, device=device) if want_shift else None,
            frac_wrapped=torch.empty((n, 3), dtype=torch.float64, device=device),
            status=torch.zeros(1, dtype=torch.int32, device=device),
        )
    else:
        if out.list.shape != (n, S) or (want_dR and out.dR is None) or (want_shift and out.shift is None):
            raise ValueError("find_neighbore_gpu: out buffers do not match")
        out.status.zero_()
    small = _lib.query("matpl_dp_neighbor_is_small_box", lat.ctypes.data, float(rc_m))
    if small or force_brute:
        dk = str(device)
        lat_d = _const(("lat", dk, lat.tobytes()), lambda: torch.as_tensor(lat.reshape(1, 9), device=device))
        fs = _const(("fs", dk, n), lambda: torch.tensor([0, n], dtype=torch.int64, device=device))
        af = _const(("af", dk, n), lambda: torch.zeros(n, dtype=torch.int32, device=device))
        _lib.call("matpl_dp_neighbor_brute", _ptr(lat_d), _ptr(frac_t), _ptr(fs), _ptr(af), _ptr(tm), _ptr(rct),
                  ntypes, m, float(rc_m), 1, n, _ptr(out.frac_wrapped), _ptr(out.list), _ptr(out.dR),
                  _ptr(out.counts), _ptr(out.shift), _ptr(out.status), _stream())
    else:
        nbytes = _lib.query("matpl_dp_neighbor_cell_workspace_bytes", lat.ctypes.data, float(rc_m), n)
        ws = torch.empty(max(int(nbytes), 16), dtype=torch.uint8, device=device)
        _lib.call("matpl_dp_neighbor_cell", lat.ctypes.data, _ptr(frac_t), _ptr(tm), _ptr(rct), ntypes, m,
                  float(rc_m), n, _ptr(out.frac_wrapped), _ptr(out.list), _ptr(out.dR), _ptr(out.counts),
                  _ptr(out.shift), _ptr(out.status), _ptr(ws), ws.numel(), _stream())
    return out.check() if check else out


def find_neighbore_batch(lattices, fracs: List, type_maps: List, ntypes: int, m: int, rc_m: float,
                         rc_type=None, want_dR=True, want_shift=False, device="cuda", check=True) -> NeighborList:
    """Many frames packed along the atom axis (brute-force kernel, any box size)."""
    lat = torch.as_tensor(np.stack([np.asarray(L, np.float64).reshape(9) for L in lattices]), device=device)
    sizes = [len(f) for f in fracs]
    start = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    ntot = int(start[-1])
    frac_t = torch.as_tensor(np.concatenate([np.asarray(f, np.float64) for f in fracs]), device=device)
    tm = torch.as_tensor(np.concatenate([np.asarray(t, np.int32) for t in type_maps]), device=device)
    af = torch.as_tensor(np.repeat(np.arange(len(sizes), dtype=np.int32), sizes), device=device)
    fs = torch.as_tensor(start, device=device)
    S = ntypes * m
    rct = _rc_type_tensor(rc_type, ntypes, rc_m, device)
    out = NeighborList(
        list=torch.empty((ntot, S), dtype=torch.int32, device=device),
        dR=torch.empty((ntot, S, 3), dtype=torch.float64, device=device) if want_dR else None,
        counts=torch.empty((ntot, ntypes), dtype=torch.int32, device=device),
        shift=torch.empty((ntot, S), dtype=torch.int32, device=device) if want_shift else None,
        frac_wrapped=torch.empty((ntot, 3), dtype=torch.float64, device=device),
        status=torch.zeros(1, dtype=torch.int32, device=device),
    )
    _lib.call("matpl_dp_neighbor_brute", _ptr(lat), _ptr(frac_t), _ptr(fs), _ptr(af), _ptr(tm), _ptr(rct), ntypes, m,
              float(rc_m), len(sizes), ntot, _ptr(out.frac_wrapped), _ptr(out.list), _ptr(out.dR), _ptr(out.counts),
              _ptr(out.shift), _ptr(out.status), _stream())
    return out.check() if check else out


def reverse_map(nl: NeighborList, type_map: torch.Tensor, m: int, frame_start=None, atom_frame=None, out=None):
    """(rev [N, T*m], status [1]): for each used entry, the index of its partner entry (j -> i,
    opposite image); status != 0 if some entry has no partner (the list is not symmetric; the fused
    force gather needs it). ``out``: (rev, status) buffers to write into.
    """
    n, S = nl.list.shape
    T = S // m
    if out is None:
        rev = torch.empty((n, S), dtype=torch.int32, device=nl.list.device)
        status = torch.zeros(1, dtype=torch.int32, device=nl.list.device)
    else:
        rev, status = out
        status.zero_()
    tm = type_map.to(torch.int32).contiguous()
    _lib.call("matpl_dp_neighbor_reverse", _ptr(nl.list), _ptr(nl.shift), _ptr(nl.counts), _ptr(tm),
              _ptr(frame_start), _ptr(atom_frame), T, m, n, _ptr(rev), _ptr(status), _stream())
    return rev, status


@dataclass
class LmpNeighData:
    imagetype_map: torch.Tensor     # [inum] int32, model type of each centre
    neighbor_list: torch.Tensor     # [1, inum, T*m] int32, j + 1 over local + ghost atoms, 0 = empty
    dR_neigh: torch.Tensor          # [1, inum, T*m, 4] float64, [r, dx, dy, dz]
    counts: torch.Tensor            # [inum, T] int32 neighbors found per bucket
    min_rsq: float                  # smallest accepted |dR|^2 (inf if none)


TOO_CLOSE_RSQ = 0.61               # pair_matpl.cpp:605-611 (compared with |dR|^2)


def generate_neighdata(x, atom_type, type_to_model, ilist, numneigh, firstneigh, neigh, ntypes: int, m: int,
                       rcut: float, check=True) -> LmpNeighData:
    """PairMATPL::generate_neighdata (src/lmps/lammps-libtorch/MATPL/pair_matpl.cpp:522-616) on the GPU.

    x [nall, 3] float64 local-then-ghost positions, atom_type [nall] LAMMPS types (1-based),
    type_to_model [ntypes_lmp] model type per LAMMPS type (-1 = absent), and LAMMPS's full
    neighbor list as CSR: ilist [inum], numneigh [inum], firstneigh [inum] (int64 offsets into
    neigh), neigh (atom indices over local + ghost). Neighbors within rcut (<=) fill the
    buckets of their model type in list order. With ``check`` (one device sync) an overflowing
    bucket raises OverflowError (the reference overruns its buffer), bad indices or types raise
    ValueError, and a pair closer than sqrt(0.61) A raises RuntimeError as the reference does.
    """
    dev = x.device
    inum = int(ilist.numel())
    nall = int(x.shape[0])
    S = ntypes * m
    imagetype = torch.empty(inum, dtype=torch.int32, device=dev)
    lst = torch.empty((1, inum, S), dtype=torch.int32, device=dev)
    dR = torch.empty((1, inum, S, 4), dtype=torch.float64, device=dev)
    counts = torch.empty((inum, ntypes), dtype=torch.int32, device=dev)
    min_rsq = torch.full((1,), 0x7FF0000000000000, dtype=torch.int64, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    xs = x.to(torch.float64).contiguous()
    ty = atom_type.to(torch.int32).contiguous()
    t2m = torch.as_tensor(type_to_model, dtype=torch.int32).to(dev).contiguous()
    il = ilist.to(torch.int32).contiguous()
    nn = numneigh.to(torch.int32).contiguous()
    fn = firstneigh.to(torch.int64).contiguous()
    ng = neigh.to(torch.int32).contiguous()
    if nall != ty.numel() or nn.numel() < inum or fn.numel() < inum:
        raise ValueError("generate_neighdata: x/type/numneigh/firstneigh sizes do not match")
    if inum:
        _lib.call("matpl_dp_lmp_neighdata", _ptr(xs), _ptr(ty), _ptr(t2m), int(t2m.numel()), _ptr(il), _ptr(nn),
                  _ptr(fn), _ptr(ng), int(ng.numel()), inum, nall, ntypes, m, float(rcut), _ptr(imagetype), _ptr(lst), _ptr(dR),
                  _ptr(counts), _ptr(min_rsq), _ptr(status), _stream())
    rmin = float("inf")
    if check and inum:
        st, mr = int(status.item()), int(min_rsq.item())
        if st > 0:
            raise OverflowError(f"generate_neighdata: atom {st - 1} has more than m={m} neighbors of one type")
        if st < 0:
            raise ValueError({-1: "generate_neighdata: ilist names an atom outside [0, inum)",
                              -2: "generate_neighdata: a neighbor index is outside [0, nall) or a list range "
                                  "outside neigh",
                              -3: "generate_neighdata: an atom type has no model type"}[st])
        rmin = float(np.array([mr], dtype=np.int64).view(np.float64)[0])
        if rmin < TOO_CLOSE_RSQ:
            raise RuntimeError(f"there are two atoms too close, min_dR_all = {rmin}")
    return LmpNeighData(imagetype, lst, dR, counts, rmin)


def image_dr(lst: torch.Tensor, dR: torch.Tensor) -> torch.Tensor:
    """ImageDR = [|dR|, dx, dy, dz], |dR| only where list > 0 (dpuni_data_loader.py:669-673)."""
    r = torch.sqrt((dR * dR).sum(-1)) * (lst > 0)
    return torch.cat([r.unsqueeze(-1), dR], -1)


def calc_stat(lst: torch.Tensor, dR: torch.Tensor, type_map: torch.Tensor, ntypes: int, rmin: float, rmax: float):
    """davg/dstd [T, 4] on the GPU (dpuni_data_loader.py:422-560)."""
    n, S = lst.shape
    davg = torch.empty((ntypes, 4), dtype=torch.float64, device=lst.device)
    dstd = torch.empty((ntypes, 4), dtype=torch.float64, device=lst.device)
    ws = torch.empty(int(_lib.query("matpl_dp_calc_stat_workspace_elems", ntypes)), dtype=torch.float64,
                     device=lst.device)
    # converted inputs stay referenced until the launch is enqueued
    lst = lst.contiguous()
    dR = dR.contiguous()
    tm = type_map.to(torch.int32).contiguous()
    _lib.call("matpl_dp_calc_stat", _ptr(lst), _ptr(dR), _ptr(tm), n, S, ntypes, float(rmin), float(rmax),
              _ptr(davg), _ptr(dstd), _ptr(ws), ws.numel(), _stream())
    return davg, dstd


def env_mat(image_dr_t: torch.Tensor, type_map: torch.Tensor, davg_full: torch.Tensor, dstd_full: torch.Tensor,
            rmin: float, rmax: float):
    """Ri [B,N,S,4], Ri_d [B,N,S,4,3] from ImageDR [B,N,S,4] (DP.calculate_Ri, dp_dp.py:217-302)."""
    if image_dr_t.dim() == 3:
        image_dr_t = image_dr_t.unsqueeze(0)
    B, N, S, _ = image_dr_t.shape
    dt = image_dr_t.dtype
    sfx = {torch.float32: "f32", torch.float64: "f64"}[dt]
    ntypes = davg_full.shape[0]
    idr = image_dr_t.contiguous()
    Ri = torch.empty((B, N, S, 4), dtype=dt, device=idr.device)
    Ri_d = torch.empty((B, N, S, 4, 3), dtype=dt, device=idr.device)
    # converted inputs stay referenced until the launch is enqueued
    tm = type_map.to(torch.int32).contiguous()
    davg = davg_full.to(dt).contiguous()
    dstd = dstd_full.to(dt).contiguous()
    _lib.call(f"matpl_dp_env_mat_{sfx}", _ptr(idr), _ptr(tm), _ptr(davg), _ptr(dstd), B, N, S, ntypes,
              float(rmin), float(rmax), _ptr(Ri), _ptr(Ri_d), _stream())
    return Ri, Ri_d
