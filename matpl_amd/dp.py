"""DP se_e2_a model on the GPU: the reference-shaped module and the fused path.

``DPTorch`` is the plain drop-in path: the same computation as the
reference's ``DP.forward`` (src/model/dp_dp.py:167-530) with the nets as
torch tensor code in the model dtype, dE/dRi by autograd, and this build's
HIP ops for everything the reference does in ``CalcOps_cuda``:
environment matrix (matpl_amd env_mat), force and virial scatter
(CalcOps_cuda.calculateForce / calculateVirial). In FP64 this is the
baseline the fused path is measured against.
"""
from __future__ import annotations

import torch

from . import calcops
from .compress import calculateCompress
from .dpnet import DPWeights
from .neighbor import env_mat
from .tallmm import add_bias, tall_mm


def _t(a, dtype, device):
    return None if a is None else torch.as_tensor(a, dtype=dtype, device=device)


class _Net(torch.nn.Module):
    """One embedding (kind 0) or fitting (kind 1) net, dp_embedding.py:40-160 layer rules."""

    def __init__(self, layers, kind: int, dtype, device):
        super().__init__()
        self.kind = kind
        self.W = torch.nn.ParameterList([torch.nn.Parameter(_t(W, dtype, device)) for W, _, _ in layers])
        self.b = torch.nn.ParameterList([torch.nn.Parameter(_t(b, dtype, device)) for _, b, _ in layers])
        self.has_rdt = [r is not None for _, _, r in layers]
        self.r = torch.nn.ParameterList([torch.nn.Parameter(_t(r if r is not None else [0.0], dtype, device))
                                         for _, _, r in layers])

    def forward(self, x, mm=None, add=None):
        """mm(x, W): the product to use (default: the plain matmul); add(y, b): the bias add."""
        nl = len(self.W)
        for l in range(nl):
            y = x @ self.W[l] if mm is None else mm(x, self.W[l])
            z = y + self.b[l] if add is None else add(y, self.b[l])
            if self.kind == 1 and l == nl - 1:
                return z
            h = torch.tanh(z)
            hs = h * self.r[l] if self.has_rdt[l] else h
            ni, no = self.W[l].shape
            if self.kind == 0:
                if no == ni:
                    x = hs + x
                elif no == 2 * ni:
                    x = torch.cat((x, x), -1) + hs
                else:
                    x = h
            else:
                x = h if l == 0 else hs + x
        return x


class DPTorch(torch.nn.Module):
    def __init__(self, w: DPWeights, dtype=torch.float64, device="cuda", compress=None, tall: bool = False):
        """compress: a matpl_amd.compress.CompressTable; the embedding nets are then evaluated from
        the table with CalcOps_cuda.calculateCompress (dp_dp.py:350-355, 572-601).
        tall: layer products and bias adds through matpl_amd.tallmm, whose weight and bias
        gradients are split over rows (for training: hipBLAS runs x^T g over all slots in one tile)."""
        super().__init__()
        cfg = w.cfg
        self.cfg = cfg
        self.dtype = dtype
        self.compress = compress
        self.mm, self.add = (tall_mm, add_bias) if tall else (None, None)
        T = cfg.ntypes
        self.emb = torch.nn.ModuleList([_Net(w.emb[p], 0, dtype, device) for p in range(T * T)])
        self.fit = torch.nn.ModuleList([_Net(w.fit[t], 1, dtype, device) for t in range(T)])
        self.register_buffer("davg_full", torch.as_tensor(w.davg_full(), dtype=dtype, device=device))
        self.register_buffer("dstd_full", torch.as_tensor(w.dstd_full(), dtype=dtype, device=device))

    def embed(self, ti: int, tj: int, Rt: torch.Tensor) -> torch.Tensor:
        """G of the (ti, tj) embedding net at the slots' normalized s = Rt[..., 0]."""
        if self.compress is None:
            return self.emb[ti * self.cfg.ntypes + tj](Rt[..., 0:1], self.mm, self.add)
        f2, coef = self.compress.lookup(Rt[..., 0], ti * self.cfg.ntypes + tj)
        return calculateCompress(f2, coef)[0].reshape(Rt.shape[:-1] + (-1,))

    def forward(self, list_neigh: torch.Tensor, Imagetype_map: torch.Tensor, ImageDR: torch.Tensor,
                nghost: int = 0, type_present=None, create_graph: bool = True):
        """list_neigh [B,N,S] 1-based; Imagetype_map [N]; ImageDR [B,N,S,4].

        type_present: the image's atom_type list as T flags (dp_dp.py:150-165, 197; None = every
        model type), or, for frames packed along the atom axis, a [N, T] bool tensor holding each
        atom's frame's flags; only the listed types' buckets take part and A is divided by m
        times their number (dp_dp.py:413).

        Returns Etot [B,1], Ei [B,N], Force [B,N+nghost,3], Virial [B,9].
        """
        cfg = self.cfg
        B, N, S = list_neigh.shape
        T, m = cfg.ntypes, cfg.m
        per_atom = isinstance(type_present, torch.Tensor) and type_present.dim() == 2
        if per_atom:
            amask = type_present.to(device=ImageDR.device, dtype=torch.bool)
            present = [t for t in range(T) if bool(amask[:, t].any())]
        else:
            present = list(range(T)) if type_present is None else [t for t in range(T) if type_present[t]]
        Ri, Ri_d = env_mat(ImageDR.to(self.dtype), Imagetype_map, self.davg_full, self.dstd_full,
                           cfg.rcut_smooth, cfg.rcut)
        Ri.requires_grad_()
        Ei = torch.zeros((B, N, 1), dtype=self.dtype, device=Ri.device)
        tmap = Imagetype_map.reshape(-1)
        for ti in present:
            idx = torch.nonzero(tmap == ti).reshape(-1)
            if idx.numel() == 0:
                continue
            A = None
            for tj in present:
                Rt = Ri[:, idx, tj * m:(tj + 1) * m]              # [B, n, m, 4]
                G = self.embed(ti, tj, Rt)                                # [B, n, m, E]
                part = Rt.transpose(-2, -1) @ G                     # [B, n, 4, E]
                if per_atom:
                    part = part * amask[idx, tj].to(self.dtype).view(1, -1, 1, 1)
                A = part if A is None else A + part
            if per_atom:
                A = A / (m * amask[idx].sum(1).to(self.dtype)).view(1, -1, 1, 1)
            else:
                A = A / (m * len(present))
            DR = A.transpose(-2, -1) @ A[..., :cfg.M2]              # [B, n, E, M2]
            Ei_t = self.fit[ti](DR.reshape(B, idx.numel(), -1), self.mm, self.add)   # [B, n, 1]
            Ei = Ei.index_copy(1, idx, Ei_t)
        Etot = Ei.sum(1)
        dE = torch.autograd.grad([Etot], [Ri], grad_outputs=[torch.ones_like(Etot)], retain_graph=True,
                                 create_graph=create_graph)[0]
        Ri_d_v = Ri_d.view(B, N, -1, 3)
        dE_v = dE.view(B, N, 1, -1)
        F = -(dE_v @ Ri_d_v).squeeze(-2)
        nl0 = (list_neigh - 1).to(torch.int32).unsqueeze(2)        # [B,N,1,S] 0-based, as dp_dp.py:466-467
        Force = calcops.calculateForce(nl0, dE_v, Ri_d_v, F, nghost)[0]
        Virial = calcops.calculateVirial(nl0, dE_v, ImageDR[..., 1:].to(self.dtype).contiguous(), Ri_d_v, nghost)[0]
        return Etot, Ei.squeeze(-1), Force, Virial
