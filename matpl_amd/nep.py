"""NEP model on the HIP operators (the GPU branch of the reference's NEP).

Mirrors src/model/nep_net.py:293-648 (device != cpu): features from
CalcOps_cuda.calculateNepFeat / calculateNepMbFeat, per-type fitting nets,
dE = dEtot/dRi by autograd (create_graph, so force losses train the
parameters), forces and virial returned negated as the reference does. The
reference's sequence per list (Ri_d from [r, x, y, z], nep_net.py:440-468;
the centre force by a batched torch.matmul, nep_net.py:608;
calculateNepForce; the frame virial of calculateNepVirial) runs here as one
kernel that forms Ri_d in registers (nep_ops.nep_force_virial), with its
backward in one kernel too; the reference's operators themselves stay
available with their own signatures (nep_ops / calcops).
Fitting-layer products run tall-by-small or split over atoms (_TallMM), at
every derivative order.

Host-side input packing (``pack_frames``) restates expand_box
(src/pre_data/nep_data_loader.py:220-254): replicate each cell num_cell times
along each vector so that the expanded box is at least 2 rc thick, keep the
inverse of the expanded box (column-major vectors) in box[9:18].
"""
from __future__ import annotations

import math
from typing import List, Optional

import numpy as np
import torch

from . import _lib, calcops, nep_ops
from .calcops import _ptr, _stream
from .tallmm import _SPLIT_ROWS, _TallMM, _tall_tn  # noqa: F401
from .nepnet import NEPWeights


def _det(l):
    return (l[0] * (l[4] * l[8] - l[5] * l[7]) - l[1] * (l[3] * l[8] - l[5] * l[6]) +
            l[2] * (l[3] * l[7] - l[4] * l[6]))


def _area(u, v):
    cx = u[1] * v[2] - u[2] * v[1]
    cy = u[2] * v[0] - u[0] * v[2]
    cz = u[0] * v[1] - u[1] * v[0]
    return math.sqrt(cx * cx + cy * cy + cz * cz)


def expand_box(lattice, cutoff):
    """(num_cell [3], box [18], box_orig [9]) for one frame; lattice rows are the cell vectors."""
    l = [float(x) for x in np.asarray(lattice, np.float64).T.reshape(9)]
    a, b, c = l[0::3], l[1::3], l[2::3]
    det = _det(l)
    vol = abs(det)
    nc = [int(math.ceil(2.0 * cutoff / (vol / _area(b, c)))), int(math.ceil(2.0 * cutoff / (vol / _area(c, a)))),
          int(math.ceil(2.0 * cutoff / (vol / _area(a, b))))]
    box = [0.0] * 18
    for k in range(3):
        for d in range(3):
            box[3 * k + d] = l[3 * k + d] * float(nc[d])
    box[9] = box[4] * box[8] - box[5] * box[7]
    box[10] = box[2] * box[7] - box[1] * box[8]
    box[11] = box[1] * box[5] - box[2] * box[4]
    box[12] = box[5] * box[6] - box[3] * box[8]
    box[13] = box[0] * box[8] - box[2] * box[6]
    box[14] = box[2] * box[3] - box[0] * box[5]
    box[15] = box[3] * box[7] - box[4] * box[6]
    box[16] = box[1] * box[6] - box[0] * box[7]
    box[17] = box[0] * box[4] - box[1] * box[3]
    det *= float(nc[0] * nc[1] * nc[2])
    for n in range(9, 18):
        box[n] /= det
    return np.array(nc, np.int64), np.array(box), np.array(l)


def pack_frames(frames, cutoff_radial, device="cuda"):
    """Packed NEP inputs (num_atoms, box, box_orig, num_cell, position, atom_type_map) as device tensors."""
    nums, boxes, origs, ncs = [], [], [], []
    for f in frames:
        nc, box, lat = expand_box(f.lattice, cutoff_radial)
        nums.append(f.natoms); boxes.append(box); origs.append(lat); ncs.append(nc)
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
    return dict(num_atoms=t(np.array(nums), torch.int64), box=t(np.concatenate(boxes), torch.float64),
                box_orig=t(np.concatenate(origs), torch.float64), num_cell=t(np.concatenate(ncs), torch.int64),
                position=t(np.concatenate([f.pos for f in frames]), torch.float64),
                atom_type_map=t(np.concatenate([f.types for f in frames]), torch.int64))


def neighbors(packed, rc_r, rc_a, ntypes, max_r=None, max_a=None, ri_dtype=None):
    """calculate_maxneigh then calculate_neighbor with Rij (as the reference's NEP data path,
    nep_data_loader.py / nep_net.py), lists sized by the largest count unless given. ri_dtype:
    torch.float32 gives the Ri rows in float (a float model's input) straight from the search.

    The list widths found by calculate_maxneigh are kept in ``packed`` (per cutoff pair): a later
    call on the same batch (a training loop revisiting it, or MD between rebuilds of the batch)
    builds its lists at those widths directly and runs the counting pass again only when they
    overflow. The lists are then at least as wide as the largest count, never narrower; the extra
    slots are empty (-1), which the descriptor and force ops skip."""
    p = packed
    key = ("_nep_list_widths", float(rc_r), float(rc_a))
    if max_r is None or max_a is None:
        widths = p.get(key)
        # (inside nep_ops.static_frames overflow is recorded, not raised: no reuse there)
        if widths is not None and nep_ops._STATIC is None:
            try:
                return _build_lists(p, rc_r, rc_a, max_r or widths[0], max_a or widths[1], ri_dtype)
            except OverflowError:
                pass
        NNr, NNa = nep_ops.calculate_maxneigh(p["num_atoms"], p["box"], p["box_orig"], p["num_cell"], p["position"],
                                              rc_r, rc_a, ntypes, p["atom_type_map"], False)
        wr = max(int(NNr.max().item()) if NNr.numel() else 1, 1)
        wa = max(int(NNa.max().item()) if NNa.numel() else 1, 1)
        if max_r is None and max_a is None:
            p[key] = (wr, wa)
        max_r, max_a = max_r or wr, max_a or wa
    return _build_lists(p, rc_r, rc_a, max_r, max_a, ri_dtype)


def _build_lists(p, rc_r, rc_a, max_r, max_a, ri_dtype=None):
    NN_r, NN_a, NL_r, NL_a, Ri_r, Ri_a = nep_ops.calculate_neighbor(
        p["num_atoms"], p["atom_type_map"], p["atom_type_map"], p["box"], p["box_orig"], p["num_cell"], p["position"],
        rc_r, rc_a, max_r, max_a, True, ri_dtype=ri_dtype)
    return dict(NN_r=NN_r, NN_a=NN_a, NL_r=NL_r, NL_a=NL_a, Ri_r=Ri_r, Ri_a=Ri_a)


def ri_d(Ri):
    """d[r, x, y, z]/d(dR) [N, S, 4, 3] (nep_net.py:440-468): row 0 = dR/r, rows 1-3 = identity."""
    N, S = Ri.shape[:2]
    out = torch.zeros((N, S, 4, 3), dtype=Ri.dtype, device=Ri.device)
    r = Ri[..., 0]
    m = (r.abs() > 1e-5).unsqueeze(-1)
    out[:, :, 0, :] = torch.where(m, Ri[..., 1:4] / torch.where(m, r.unsqueeze(-1), torch.ones_like(r).unsqueeze(-1)),
                                  torch.zeros_like(Ri[..., 1:4]))
    eye = torch.eye(3, dtype=Ri.dtype, device=Ri.device)
    out[:, :, 1:4, :] = eye * m.unsqueeze(-1).to(Ri.dtype)
    return out


class _FitNet(torch.nn.Module):
    """NEP fitting net (src/model/nep_fitting.py:40-163): tanh first layer, residual hidden layers
    (with resnet_dt), linear output."""

    def __init__(self, layers, dtype, device):
        super().__init__()
        self.W = torch.nn.ParameterList()
        self.b = torch.nn.ParameterList()
        self.rdt = []
        for W, b, r in layers:
            self.W.append(torch.nn.Parameter(torch.as_tensor(W, dtype=dtype, device=device)))
            self.b.append(torch.nn.Parameter(torch.as_tensor(b, dtype=dtype, device=device)))
            self.rdt.append(None if r is None else torch.nn.Parameter(torch.as_tensor(r, dtype=dtype, device=device)))
        for k, r in enumerate(self.rdt):
            if r is not None:
                self.register_parameter(f"rdt{k}", r)

    def forward(self, x):
        L = len(self.W)
        for l in range(L):
            z = _TallMM.apply(x, self.W[l], False) + self.b[l]
            if l == L - 1:
                return z
            h = torch.tanh(z)
            x = h if l == 0 else (h * self.rdt[l] if self.rdt[l] is not None else h) + x
        return x


class FitPlan:
    """Atoms grouped by type and cut into tiles for the fitting kernels (nep_fit.hip): perm [N]
    (int32, the atoms of type 0, then type 1, ...), tiles [ntiles, 3] (type, first position in
    perm, count), tstart [T + 1] (type t's tiles)."""

    def __init__(self, type_index, f64: bool, D: int, H: int, device):
        T = len(type_index)
        ta = _lib.query("matpl_nep_fit_tile_atoms", int(f64), D, H)
        if ta < 0:
            raise ValueError(f"NEP fitting kernel: no tile fits D={D}, H={H}")
        counts = [int(ix.numel()) for ix in type_index]
        rows, tstart, off = [], [0], 0
        for t, c in enumerate(counts):
            for s0 in range(0, c, ta):
                rows.append((t, off + s0, min(ta, c - s0)))
            off += c
            tstart.append(len(rows))
        self.tile_atoms, self.ntiles, self.ntypes = ta, len(rows), T
        self.perm = (torch.cat(list(type_index)).to(torch.int32) if T else
                     torch.zeros(0, dtype=torch.int32, device=device)).contiguous()
        self.tiles = torch.as_tensor(np.asarray(rows, np.int32).reshape(-1, 3), device=device)
        self.tstart = torch.as_tensor(np.asarray(tstart, np.int32), device=device)


class _NepFitFn(torch.autograd.Function):
    """(Ei [N], g = dEi/dq [N, D]) of the per-type one-hidden-layer nets; the backward takes
    dL/dEi and dL/dg and returns dL/dq and the parameter gradients (nep_fit.hip)."""

    @staticmethod
    def forward(ctx, q, qs, W0, b0, w1, b1, plan):
        q = q.contiguous()
        N, D = q.shape
        H = W0.shape[2]
        Ei = torch.empty(N, dtype=q.dtype, device=q.device)
        g = torch.empty_like(q)
        _lib.call("matpl_nep_fit_forward", int(q.dtype == torch.float64), _ptr(q), _ptr(qs), _ptr(plan.perm),
                  _ptr(plan.tiles), plan.ntiles, plan.tile_atoms, _ptr(W0), _ptr(b0), _ptr(w1), _ptr(b1), D, H,
                  _ptr(Ei), _ptr(g), _stream())
        ctx.save_for_backward(q, qs, W0, b0, w1, b1)
        ctx.plan = plan
        return Ei, g

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gE, gg):
        q, qs, W0, b0, w1, b1 = ctx.saved_tensors
        plan = ctx.plan
        N, D = q.shape
        H = W0.shape[2]
        f64 = int(q.dtype == torch.float64)
        gE = None if gE is None else gE.contiguous()
        gg = None if gg is None else gg.contiguous()
        dq = torch.empty_like(q) if ctx.needs_input_grad[0] else None
        part = torch.empty(max(plan.ntiles, 1) * _lib.query("matpl_nep_fit_part_width", D, H), dtype=q.dtype,
                           device=q.device)
        dW0, db0, dw1, db1 = (torch.empty_like(x) for x in (W0, b0, w1, b1))
        _lib.call("matpl_nep_fit_backward", f64, _ptr(q), _ptr(qs), _ptr(plan.perm), _ptr(plan.tiles), plan.ntiles,
                  plan.tile_atoms, _ptr(plan.tstart), plan.ntypes, _ptr(W0), _ptr(b0), _ptr(w1), _ptr(b1), D, H,
                  _ptr(gE), _ptr(gg), _ptr(dq), _ptr(part), _ptr(dW0), _ptr(db0), _ptr(dw1), _ptr(db1), _stream())
        return dq, None, dW0, db0, dw1, db1, None


class NEPTorch(torch.nn.Module):
    """NEP.forward on the HIP operators: (Etot [F,1], Ei [N], Force [N,3], Virial [F,9])."""

    def __init__(self, w: NEPWeights, dtype=torch.float64, device="cuda"):
        super().__init__()
        calcops.register()
        cfg = w.cfg
        self.cfg = cfg
        self.c2 = torch.nn.Parameter(torch.as_tensor(w.c2, dtype=dtype, device=device))
        self.c3 = torch.nn.Parameter(torch.as_tensor(w.c3, dtype=dtype, device=device))
        self.register_buffer("q_scaler", torch.as_tensor(w.q_scaler, dtype=dtype, device=device))
        self.fit = torch.nn.ModuleList([_FitNet(layers, dtype, device) for layers in w.fit])
        self.dtype = dtype
        # one hidden layer with biases in every net (the NEP form, network_size [H, 1]): the fitting
        # kernels; otherwise the per-type torch layers
        self.fit_kernel = all(len(n.W) == 2 and n.W[1].shape[1] == 1 and all(r is None for r in n.rdt)
                              for n in self.fit) and len({tuple(n.W[0].shape) for n in self.fit}) == 1
        self._plans = {}

    def features(self, Ri_r, NL_r, Ri_a, NL_a, atom_type_map):
        cfg = self.cfg
        l3, l4, l5 = cfg.l_max
        ops = torch.ops.CalcOps_cuda
        q = ops.calculateNepFeat(self.c2, Ri_r, NL_r, atom_type_map, Ri_r.new_empty(0), cfg.rcut_radial,
                                 cfg.n_angular_feats)[0]
        if l3 > 0:
            qa = ops.calculateNepMbFeat(self.c3, Ri_a, NL_a, atom_type_map, Ri_a.new_empty(0), cfg.nr, l3, l4, l5,
                                        cfg.rcut_angular)[0]
            q = torch.cat([q, qa], -1)
        return q

    def fit_plan(self, atom_type_map, type_index=None) -> FitPlan:
        """The fitting kernels' tiles; cached for a given ``type_index`` list (graph replays)."""
        D, H = self.fit[0].W[0].shape
        f64 = self.dtype == torch.float64
        if type_index is not None:
            key = tuple(ix.data_ptr() for ix in type_index) + tuple(ix.numel() for ix in type_index)
            plan = self._plans.get(key)
            if plan is None:
                plan = FitPlan(type_index, f64, D, H, atom_type_map.device)
                self._plans = {key: plan}
            return plan
        # one plan per type map (and its version): the per-type counts are read on the host once
        return nep_ops._cached(atom_type_map, ("fit_plan", f64, D, H, self.cfg.ntypes),
                               lambda: self._plan_of(atom_type_map, f64, D, H))

    def _plan_of(self, atom_type_map, f64, D, H):
        tm = atom_type_map.reshape(-1)
        perm = torch.argsort(tm, stable=True)
        counts = torch.bincount(tm, minlength=self.cfg.ntypes).cpu().tolist()
        return FitPlan(list(torch.split(perm, counts)), f64, D, H, tm.device)

    def energies_and_grad(self, feats, atom_type_map, type_index=None):
        """(Ei [N], dEi/dfeats [N, D]) from the fitting kernels (fit_kernel nets only)."""
        plan = self.fit_plan(atom_type_map, type_index)
        W0 = torch.stack([n.W[0] for n in self.fit])
        b0 = torch.stack([n.b[0] for n in self.fit])
        w1 = torch.stack([n.W[1][:, 0] for n in self.fit])
        b1 = torch.stack([n.b[1].reshape(()) for n in self.fit])
        return _NepFitFn.apply(feats, self.q_scaler, W0, b0, w1, b1, plan)

    def energies(self, feats, atom_type_map, type_index=None):
        """type_index: the atoms of each type (precomputed, e.g. for graph capture); else found here."""
        if self.fit_kernel:
            return self.energies_and_grad(feats, atom_type_map, type_index)[0]
        x = feats * self.q_scaler
        Ei = torch.zeros(feats.shape[0], dtype=feats.dtype, device=feats.device)
        for t, net in enumerate(self.fit):
            idx = torch.nonzero(atom_type_map == t).reshape(-1) if type_index is None else type_index[t]
            if idx.numel():
                Ei = Ei.index_put((idx,), net(x[idx]).reshape(-1))
        return Ei

    def forward(self, NN_r, NL_r, Ri_r, NN_a, NL_a, Ri_a, num_atom, atom_type_map, is_calc_f=True,
                create_graph=None, type_index=None):
        Ri_r = Ri_r.to(self.dtype).detach().requires_grad_()
        Ri_a = Ri_a.to(self.dtype).detach().requires_grad_()
        l3 = self.cfg.l_max[0]
        feats = self.features(Ri_r, NL_r, Ri_a, NL_a, atom_type_map)
        gq = None
        if self.fit_kernel:
            Ei, gq = self.energies_and_grad(feats, atom_type_map, type_index)
        else:
            Ei = self.energies(feats, atom_type_map, type_index)
        nf = num_atom.numel()
        frame = nep_ops._frame_index(num_atom, Ei.device)[1]
        if nf == 1:
            # one frame: a plain reduction (index_add would send every atom's atomic add to one address)
            Etot = Ei.sum().reshape(1, 1)
        else:
            Etot = torch.zeros(nf, dtype=Ei.dtype, device=Ei.device).index_add(0, frame, Ei).unsqueeze(-1)
        if not is_calc_f:
            return Etot, Ei, None, None
        if create_graph is None:
            create_graph = torch.is_grad_enabled() and any(p.requires_grad for p in self.parameters())
        ins = [Ri_r, Ri_a] if l3 > 0 else [Ri_r]
        if gq is not None:
            # dEtot/dRi = (dEi/dq) dq/dRi: the fitting kernel's g through the feature ops only
            g = torch.autograd.grad([feats], ins, grad_outputs=[gq], create_graph=create_graph, allow_unused=True)
        else:
            g = torch.autograd.grad([Etot.sum()], ins, create_graph=create_graph, allow_unused=True)
        Force = Virial = None
        for Ri, NL, dE in zip(ins, (NL_r, NL_a), g):
            if dE is None:
                continue
            # ri_d, the centre term, calculateNepForce and the frame sum of calculateNepVirial in one
            # pass, with the position derivative formed from Ri in the kernel
            F, V = nep_ops.nep_force_virial(NL, dE, Ri, num_atom)
            Force = F if Force is None else Force + F
            Virial = V if Virial is None else Virial + V
        return Etot, Ei, -Force, -Virial


class NEPGraph:
    """E+F+V of a fixed set of frames, neighbor search included, replayed as one HIP graph.

    The positions are read from ``packed["position"]`` at every replay (update it in place). The
    per-type atom lists and frame index are fixed at construction; the neighbor lists are sized by
    ``max_r`` / ``max_a``, and overflow is reported by ``check()``. An eager call first runs the
    checked ops on the same inputs; the captured sequence makes no host synchronisation
    (nep_ops.static_frames)."""

    def __init__(self, model: NEPTorch, packed, ntypes: int, max_r: int, max_a: int):
        self.model, self.p = model, packed
        cfg = model.cfg
        self.args = (cfg.rcut_radial, cfg.rcut_angular, ntypes, max_r, max_a)
        tm = packed["atom_type_map"]
        self.type_index = [torch.nonzero(tm == t).reshape(-1) for t in range(ntypes)]
        self.eager_out = self._run()                     # checked ops on these inputs
        cur = torch.cuda.current_stream()
        side = torch.cuda.Stream()
        side.wait_stream(cur)
        with torch.cuda.stream(side):
            with nep_ops.static_frames(packed["num_atoms"]):
                for _ in range(2):
                    self._run()
        cur.wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with nep_ops.static_frames(packed["num_atoms"]) as st:
            with torch.cuda.graph(self.graph):
                self.out = self._run()
        self.frames = st                                 # keeps the frame index and status alive

    def _run(self):
        p = self.p
        # a float model takes float rows straight from the search
        nb = neighbors(p, *self.args, ri_dtype=torch.float32 if self.model.dtype == torch.float32 else None)
        return self.model(nb["NN_r"], nb["NL_r"], nb["Ri_r"], nb["NN_a"], nb["NL_a"], nb["Ri_a"], p["num_atoms"],
                          p["atom_type_map"], create_graph=False, type_index=self.type_index)

    def replay(self):
        """(Etot, Ei, Force, Virial) at the current positions (tensors reused by every replay)."""
        self.graph.replay()
        return self.out

    def check(self):
        s = int(self.frames.status.item())
        if s:
            raise OverflowError(f"NEPGraph: more than max_NN neighbors (atom {s - 1})")
