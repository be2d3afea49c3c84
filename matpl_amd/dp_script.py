"""A TorchScript DP model for hosts without Python (the LAMMPS pair style, pair_matpl.cpp:864-1000).

``DPScript`` computes DP.forward (src/model/dp_dp.py:167-530) with the call signature the pair style
uses for a scripted model (pair_matpl.cpp:925-931):

    forward(list_neigh [1, N, T*m] int32 (j + 1 over local + ghost atoms, 0 empty),
            Imagetype_map [N], atom_type [T], ImageDR [1, N, T*m, 4] float64, nghost: int)
      -> (Etot [1, 1], Ei [1, N], Force [1, N + nghost, 3], Egroup (None), Virial [1, 9])

Operators: torch.ops.matpl_amd.env_mat (Ri, Ri_d on the HIP kernel) and
torch.ops.CalcOps_cuda.calculateForce / calculateVirial, both registered in C++ by
libCalcOps_bind.so (``load_ops``); nets in torch. Every model type counts in the 1/(m T)
normalisation of A (dp_dp.py:197, 413 with type_nums from atom_type), as in the pair style's call,
which passes all of the model's types in atom_type. Script it with ``torch.jit.script`` after
``load_ops()`` in a process that has not registered the Python CalcOps_cuda ops.
"""
from __future__ import annotations

import os
from typing import List, Optional, Tuple

import numpy as np
import torch

from .dpnet import DPWeights

OPS_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libCalcOps_bind.so")


def load_ops(path: str = OPS_LIB) -> None:
    """Registers the C++ operators (CalcOps_cuda DP ops and matpl_amd::env_mat)."""
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (python -m matpl_amd.build --torch-ops)")
    torch.ops.load_library(path)


class _Layer(torch.nn.Module):
    def __init__(self, W, b, r, mode: int, last: bool, dtype, device):
        super().__init__()
        t = lambda a: torch.as_tensor(np.asarray(a), dtype=dtype, device=device)
        self.W = torch.nn.Parameter(t(W))
        self.b = torch.nn.Parameter(t(b if b is not None else np.zeros(np.asarray(W).shape[1])))
        self.r = torch.nn.Parameter(t(r if r is not None else np.ones(np.asarray(W).shape[1])))
        self.mode = mode        # 0: x = h, 1: x = h r + x, 2: x = cat(x, x) + h r
        self.last = last        # fitting output layer: linear

    def forward(self, x):
        z = x @ self.W + self.b
        if self.last:
            return z
        h = torch.tanh(z)
        if self.mode == 1:
            return h * self.r + x
        if self.mode == 2:
            return torch.cat((x, x), -1) + h * self.r
        return h


def _net(layers, kind: int, dtype, device) -> torch.nn.ModuleList:
    out = torch.nn.ModuleList()
    n = len(layers)
    for l, (W, b, r) in enumerate(layers):
        ni, no = np.asarray(W).shape
        if kind == 0:
            mode = 1 if no == ni else (2 if no == 2 * ni else 0)
        else:
            mode = 0 if l == 0 else 1
        out.append(_Layer(W, b, r, mode, kind == 1 and l == n - 1, dtype, device))
    return out


class _Net(torch.nn.Module):
    def __init__(self, layers, kind: int, dtype, device):
        super().__init__()
        self.layers = _net(layers, kind, dtype, device)

    def forward(self, x):
        for layer in self.layers:
            x = layer(x)
        return x


class DPScript(torch.nn.Module):
    def __init__(self, w: DPWeights, dtype=torch.float64, device="cuda"):
        super().__init__()
        cfg = w.cfg
        self.ntypes = cfg.ntypes
        self.m = cfg.m
        self.M2 = cfg.M2
        self.rmin = float(cfg.rcut_smooth)
        self.rmax = float(cfg.rcut)
        self.emb = torch.nn.ModuleList([_Net(w.emb[p], 0, dtype, device) for p in range(cfg.ntypes ** 2)])
        self.fit = torch.nn.ModuleList([_Net(w.fit[t], 1, dtype, device) for t in range(cfg.ntypes)])
        self.register_buffer("davg_full", torch.as_tensor(w.davg_full(), dtype=dtype, device=device))
        self.register_buffer("dstd_full", torch.as_tensor(w.dstd_full(), dtype=dtype, device=device))

    def forward(self, list_neigh: torch.Tensor, Imagetype_map: torch.Tensor, atom_type: torch.Tensor,
                ImageDR: torch.Tensor, nghost: int) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor,
                                                            Optional[torch.Tensor], torch.Tensor]:
        B, N, S = list_neigh.shape[0], list_neigh.shape[1], list_neigh.shape[2]
        T, m = self.ntypes, self.m
        dt = self.davg_full.dtype
        tmap = Imagetype_map.reshape(-1).to(torch.int64)
        out = torch.ops.matpl_amd.env_mat(ImageDR.to(dt), tmap, self.davg_full, self.dstd_full, self.rmin, self.rmax)
        Ri = out[0].detach().requires_grad_(True)
        Ri_d = out[1]
        idx_list: List[torch.Tensor] = []
        A_list: List[Optional[torch.Tensor]] = []
        for t in range(T):
            idx_list.append(torch.nonzero(tmap == t).reshape(-1))
            A_list.append(None)
        for p, net in enumerate(self.emb):
            ti = p // T
            tj = p % T
            idx = idx_list[ti]
            if idx.numel() > 0:
                Rt = Ri[:, idx, tj * m:(tj + 1) * m]
                part = Rt.transpose(-2, -1) @ net(Rt[..., 0:1])
                a = A_list[ti]
                A_list[ti] = part if a is None else a + part
        Ei = torch.zeros((B, N, 1), dtype=dt, device=Ri.device)
        for t, net in enumerate(self.fit):
            a = A_list[t]
            if a is not None:
                A = a / (m * T)
                DR = A.transpose(-2, -1) @ A[..., :self.M2]
                Ei = Ei.index_copy(1, idx_list[t], net(DR.reshape(B, idx_list[t].numel(), -1)))
        Etot = Ei.sum(1)
        grad_out: List[Optional[torch.Tensor]] = [torch.ones_like(Etot)]
        g = torch.autograd.grad([Etot], [Ri], grad_outputs=grad_out, retain_graph=True)[0]
        assert g is not None
        dE = g.view(B, N, 1, -1)
        Rid = Ri_d.view(B, N, -1, 3)
        F = -torch.matmul(dE, Rid).squeeze(-2)
        nl0 = (list_neigh - 1).to(torch.int32).unsqueeze(2)
        ng = torch.tensor(nghost, dtype=torch.int64)
        Force = torch.ops.CalcOps_cuda.calculateForce(nl0, dE, Rid, F, ng)[0]
        Virial = torch.ops.CalcOps_cuda.calculateVirial(nl0, dE, ImageDR[..., 1:].to(dt).contiguous(), Rid, ng)[0]
        return Etot, Ei.squeeze(-1), Force, None, Virial
