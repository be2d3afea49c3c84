"""Batched DP training step on the HIP operators: many frames of uneven size in one pass,
energy + force loss differentiated to every net parameter (force-loss double backward).

The reference trains DP (src/PWMLFF/dp_mods/dp_trainer.py:34-318) through DP.forward
(src/model/dp_dp.py:167-530): every one of the T*m slots of every atom, empty ones included,
goes through the embedding nets; dE/dRi comes from autograd with create_graph; forces from
CalcOps.calculateForce (centre term by torch.matmul, dp_dp.py:436); loss.backward()
differentiates through all of it. ``DPTorch`` in FP64 is that path here (the plain drop-in).

``DPTrain`` evaluates the same model and the same loss gradient with:
  * all frames at once: the neighbor lists of every frame are packed along one atom axis with
    global indices, so each op runs once per batch instead of once per frame;
  * each type bucket cut to its first mc slots, mc = the largest in-cutoff count of any bucket in
    the batch. The m - mc slots cut away are empty, all holding the empty-slot descriptor
    Ri_pad = -davg/dstd, and are added back in closed form: (m - mc) Ri_pad^T G(Ri_pad[0]) per
    bucket (the same sum the reference forms slot by slot);
  * descriptors Ri, Ri_d in FP32 from the HIP env-mat kernel;
  * embedding and fitting layer products on BF16-rounded operands with FP32 VALU FMA
    accumulation and FP32 activations ("bf16"), or all in FP32 ("fp32");
  * the per-atom contractions A = Ri^T G and DR = A^T A' (and all their derivatives) on the
    batched small-product kernels (matpl_amd.bsmm);
  * the centre force by a HIP reduction; neighbor forces by CalcOps_cuda.calculateForce, whose
    backward with respect to dE is the hand-written force-gradient kernel, the edge through
    which the force loss reaches the nets;
  * every layer's weight and bias gradient split over rows (matpl_amd.tallmm) at every
    derivative order.
The per-frame virial is not part of this loss (the packed batch's virial op sums all frames).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List

import numpy as np
import torch

from . import calcops, nep_ops
from .bsmm import bmm_tn
from .dp import DPTorch, _Net
from .dpnet import DPWeights
from .neighbor import env_mat, find_neighbore_batch, image_dr
from .tallmm import add_bias, tall_mm_bf16, tall_mm_f32


@dataclass
class TrainBatch:
    nframes: int
    natoms: int
    type_map: torch.Tensor       # [N] int64
    frame: torch.Tensor          # [N] int64, frame of each atom
    frame_natoms: torch.Tensor   # [F] float64
    by_type: List[torch.Tensor]  # atoms of each type
    mc: int                      # slots kept per type bucket
    list_c: torch.Tensor         # [1, N, T*mc] int32, 1-based over the packed batch, 0 = empty
    image_dr_c: torch.Tensor     # [1, N, T*mc, 4] float64
    E_label: torch.Tensor        # [F] float64
    F_label: torch.Tensor        # [N, 3] float64
    list_full: torch.Tensor = None       # [1, N, T*m] (kept for the plain path)
    image_dr_full: torch.Tensor = None
    type_present: torch.Tensor = None    # [F, T] bool: each frame's atom_type list (dp_dp.py:150-165)

    def atom_type_present(self) -> torch.Tensor:
        """[N, T] bool: the type list of each atom's frame."""
        return self.type_present[self.frame]


def pack_batch(frames, cfg, device="cuda", seed=0, keep_full=False, type_present=None) -> TrainBatch:
    """Neighbor lists of all frames (one batched launch), packed with global indices, and
    seeded synthetic labels (energies ~ N(0, 0.1) eV/atom, forces ~ N(0, 0.5) eV/A).
    type_present [F, T]: each frame's atom_type list; default the types of its atoms, as the
    reference's loader gives it per image."""
    T, m = cfg.ntypes, cfg.m
    nl = find_neighbore_batch([f.lattice for f in frames], [f.frac() for f in frames], [f.types for f in frames],
                              T, m, cfg.rcut, device=device)
    sizes = [f.natoms for f in frames]
    N = sum(sizes)
    start = torch.as_tensor(np.concatenate([[0], np.cumsum(sizes)[:-1]]), dtype=torch.int32, device=device)
    frame = torch.repeat_interleave(torch.arange(len(frames), device=device), torch.as_tensor(sizes, device=device))
    lst = nl.list.view(N, T, m)
    glob = torch.where(lst > 0, lst + start[frame].view(N, 1, 1), torch.zeros_like(lst))
    idr = image_dr(nl.list, nl.dR).view(N, T, m, 4)
    mc = max(1, int(nl.counts.max().item()))
    tm = torch.as_tensor(np.concatenate([f.types for f in frames]).astype(np.int64), device=device)
    rng = np.random.default_rng(seed)
    E_label = torch.as_tensor(rng.normal(0.0, 0.1, len(frames)) * np.asarray(sizes), device=device)
    F_label = torch.as_tensor(rng.normal(0.0, 0.5, (N, 3)), device=device)
    b = TrainBatch(len(frames), N, tm, frame, torch.as_tensor(sizes, dtype=torch.float64, device=device),
                   [torch.nonzero(tm == t).reshape(-1) for t in range(T)], mc,
                   glob[:, :, :mc].reshape(1, N, T * mc).contiguous(),
                   idr[:, :, :mc].reshape(1, N, T * mc, 4).contiguous(), E_label, F_label)
    if type_present is None:
        tp = np.zeros((len(frames), T), bool)
        for k, f in enumerate(frames):
            tp[k, np.unique(f.types)] = True
    else:
        tp = np.asarray(type_present, bool).reshape(len(frames), T)
        for k, f in enumerate(frames):
            if not tp[k, np.unique(f.types)].all():
                raise ValueError(f"type_present: frame {k} has atoms of a type outside its list")
    b.type_present = torch.as_tensor(tp, device=device)
    if keep_full:
        b.list_full = glob.reshape(1, N, T * m).contiguous()
        b.image_dr_full = idr.reshape(1, N, T * m, 4).contiguous()
    return b


def loss_fn(Etot, F, batch: TrainBatch, pref_e=1.0, pref_f=1.0):
    """MSE of the per-atom energy of each frame plus MSE of the force components
    (the E and F terms of dp_trainer.py's loss)."""
    le = (((Etot - batch.E_label) / batch.frame_natoms) ** 2).mean()
    lf = ((F - batch.F_label) ** 2).mean()
    return pref_e * le + pref_f * lf


class DPTrain(torch.nn.Module):
    def __init__(self, w: DPWeights, precision: str = "bf16", device="cuda"):
        super().__init__()
        if precision not in ("bf16", "fp32"):
            raise ValueError("precision must be 'bf16' or 'fp32'")
        cfg = w.cfg
        self.cfg = cfg
        self.mm = tall_mm_bf16 if precision == "bf16" else tall_mm_f32
        T = cfg.ntypes
        f32 = torch.float32
        self.emb = torch.nn.ModuleList([_Net(w.emb[p], 0, f32, device) for p in range(T * T)])
        self.fit = torch.nn.ModuleList([_Net(w.fit[t], 1, f32, device) for t in range(T)])
        self.register_buffer("davg", torch.as_tensor(w.davg, dtype=f32, device=device))
        self.register_buffer("dstd", torch.as_tensor(w.dstd, dtype=f32, device=device))
        self.register_buffer("ri_pad", torch.as_tensor(-w.davg / w.dstd, dtype=f32, device=device))

    def forward(self, batch: TrainBatch, create_graph: bool = True):
        """(Etot per frame [F], Ei [N], Force [N, 3]) in FP32."""
        cfg = self.cfg
        T, m, mc, N = cfg.ntypes, cfg.m, batch.mc, batch.natoms
        Sc = T * mc
        davg_c = self.davg[:, None, :].expand(T, Sc, 4).reshape(T, -1)
        dstd_c = self.dstd[:, None, :].expand(T, Sc, 4).reshape(T, -1)
        Ri, Ri_d = env_mat(batch.image_dr_c.to(torch.float32), batch.type_map, davg_c, dstd_c, cfg.rcut_smooth,
                           cfg.rcut)
        Ri.requires_grad_()
        # each frame's type list (dp_dp.py:150-165): a bucket takes part only for atoms whose frame
        # lists its type, and A is divided by m times the frame's number of listed types (:413)
        tp = batch.type_present
        listed = [t for t in range(T) if bool(tp[:, t].any())]
        every = [bool(tp[:, t].all()) for t in range(T)]
        amask = batch.atom_type_present().to(torch.float32)                    # [N, T]
        inv = 1.0 / (m * amask.sum(1))                                           # [N]
        Ei = torch.zeros(N, dtype=torch.float32, device=Ri.device)
        for ti in range(T):
            idx = batch.by_type[ti]
            if idx.numel() == 0:
                continue
            A = None
            for tj in listed:
                net = self.emb[ti * T + tj]
                Rt = Ri[0, idx, tj * mc:(tj + 1) * mc]                      # [n, mc, 4]
                G = net(Rt[..., 0:1], self.mm, add_bias)                             # [n, mc, E]
                part = bmm_tn(Rt, G)                                         # [n, 4, E]
                if m > mc:
                    gpad = net(self.ri_pad[ti, 0].view(1, 1), self.mm)        # [1, E]
                    part = part + (m - mc) * (self.ri_pad[ti].view(4, 1) * gpad)
                if not every[tj]:
                    part = part * amask[idx, tj].view(-1, 1, 1)
                A = part if A is None else A + part
            A = A * inv[idx].view(-1, 1, 1)
            DR = bmm_tn(A, A[..., :cfg.M2])
            Ei = Ei.index_copy(0, idx, self.fit[ti](DR.reshape(idx.numel(), -1), self.mm, add_bias).reshape(-1))
        Etot = torch.zeros(batch.nframes, dtype=Ei.dtype, device=Ei.device).index_add(0, batch.frame, Ei)
        dE = torch.autograd.grad([Etot.sum()], [Ri], create_graph=create_graph)[0]
        dE_v = dE.view(1, N, 1, Sc * 4)
        Ri_d_v = Ri_d.view(1, N, Sc * 4, 3)
        centre = nep_ops.nep_centre_force(dE.view(N, Sc, 4), Ri_d.view(N, Sc, 4, 3)).view(1, N, 3)
        nl0 = (batch.list_c - 1).to(torch.int32).unsqueeze(2)
        Force = calcops.calculateForce(nl0, dE_v, Ri_d_v, centre, 0)[0]
        return Etot, Ei, Force[0]


def train_grads(model: DPTrain, batch: TrainBatch):
    """One training step's parameter gradients (and the loss)."""
    params = [p for p in model.parameters() if p.requires_grad]
    Etot, Ei, F = model(batch)
    loss = loss_fn(Etot.double(), F.double(), batch)
    return loss, torch.autograd.grad(loss, params, allow_unused=True)


def plain_grads(model: DPTorch, batch: TrainBatch, chunk_frames=None):
    """The same loss and gradients through the plain FP64 path (all T*m slots), optionally
    accumulated over chunks of frames to bound memory."""
    params = [p for p in model.parameters() if p.requires_grad]
    F_all = batch.nframes
    chunks = [(0, F_all)] if chunk_frames is None else [(a, min(F_all, a + chunk_frames))
                                                        for a in range(0, F_all, chunk_frames)]
    total = None
    grads = None
    fstart = torch.cat([torch.zeros(1, dtype=torch.int64, device=batch.frame.device),
                        torch.cumsum(batch.frame_natoms.to(torch.int64), 0)]).tolist()
    for a, b in chunks:
        lo, hi = fstart[a], fstart[b]
        lst = batch.list_full[:, lo:hi]
        lst = torch.where(lst > 0, lst - lo, lst)
        tm = batch.type_map[lo:hi]
        Etot_atoms, Ei, F, V = model(lst, tm, batch.image_dr_full[:, lo:hi],
                                     type_present=batch.atom_type_present()[lo:hi], create_graph=True)
        Etot = torch.zeros(b - a, dtype=Ei.dtype, device=Ei.device).index_add(0, batch.frame[lo:hi] - a, Ei[0])
        le = (((Etot - batch.E_label[a:b]) / batch.frame_natoms[a:b]) ** 2).sum() / F_all
        lf = ((F[0] - batch.F_label[lo:hi]) ** 2).sum() / (3 * batch.natoms)
        loss = le + lf
        g = torch.autograd.grad(loss, params, allow_unused=True)
        total = loss.detach() if total is None else total + loss.detach()
        grads = list(g) if grads is None else [x if y is None else (y if x is None else x + y)
                                               for x, y in zip(grads, g)]
    return total, grads
