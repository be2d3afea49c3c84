"""DP model compression: tabulated embedding nets and the calculateCompress operator.

Table (src/user/ckpt_compress.py:20-89, 110-226, 348-377): for every type
pair the embedding net G(s) (s = normalized s_ij = Ri[..., 0]) is replaced by
piecewise polynomials on a grid of knots x_k: Hermite interpolation of G and
G' (order 3, cubic) or G, G', G'' (order 5, quintic) at both ends of each
interval, coefficients stored highest power first. Knots run from s_min
(the most negative normalized s of an empty slot, (0 - davg)/dstd) in steps
of dx up to the first knot >= s_max = max (s_max_raw - davg)/dstd, then in
steps of 10 dx up to 10 s_max.

Lookup (src/model/dp_dp.py:572-601): interval k of s and f2 = s - x_k, then
G = calculateCompress(f2, table[pair, k]) on the GPU. Unlike the reference,
knots are x_0 + k dx (not accumulated sums) and values beyond s_max index
the coarse part of the table where it lies (the reference indexes the fine
part's first rows there and offsets from s_max rather than from the knot).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .calcops import _ptr, _stream


# ------------------------------------------------------------------ operator
def _check(f2, coef):
    if coef.dim() != 3 or coef.shape[2] not in (4, 6):
        raise ValueError(f"calculateCompress: coefficient must be [n, L, 4|6], got {tuple(coef.shape)}")
    if f2.numel() != coef.shape[0]:
        raise ValueError(f"calculateCompress: f2 has {f2.numel()} entries for {coef.shape[0]} coefficient rows")
    if f2.dtype not in (torch.float32, torch.float64):
        raise RuntimeError(f"calculateCompress: unsupported dtype {f2.dtype}")


class CalculateCompress(torch.autograd.Function):
    """G [n, L] from f2 [n, 1] and coefficient [n, L, K] (CalcOps.cpp:331-383); grad wrt f2 only."""

    @staticmethod
    def forward(ctx, f2, coef):
        _check(f2, coef)
        f2c = f2.contiguous()
        cc = coef.to(f2.dtype).contiguous()
        n, L, K = cc.shape
        G = torch.empty((n, L), dtype=f2.dtype, device=f2.device)
        name = "matpl_dp_compress_f64" if f2.dtype == torch.float64 else "matpl_dp_compress_f32"
        _lib.call(name, _ptr(f2c), _ptr(cc), n, L, K, _ptr(G), _stream())
        ctx.save_for_backward(f2c, cc)
        ctx.f2_shape = f2.shape
        return G

    @staticmethod
    def backward(ctx, gG):
        f2, cc = ctx.saved_tensors
        n, L, K = cc.shape
        gG = gG.to(f2.dtype).contiguous()
        g = torch.empty(n, dtype=f2.dtype, device=f2.device)
        name = "matpl_dp_compress_grad_f64" if f2.dtype == torch.float64 else "matpl_dp_compress_grad_f32"
        _lib.call(name, _ptr(f2), _ptr(cc), _ptr(gG), n, L, K, _ptr(g), _stream())
        return g.reshape(ctx.f2_shape), None


def calculateCompress(f2, coefficient):
    return [CalculateCompress.apply(f2, coefficient)]


# ------------------------------------------------------------------ table
@dataclass
class CompressTable:
    table: torch.Tensor      # [T*T, rows, E, K]
    x0: float                # first knot (s_min)
    dx: float
    n_fine: int              # intervals of width dx
    n_coarse: int            # intervals of width 10 dx after them
    order: int

    @property
    def x_fine_end(self) -> float:
        return self.x0 + self.n_fine * self.dx

    def knots(self) -> np.ndarray:
        fine = self.x0 + self.dx * np.arange(self.n_fine + 1)
        coarse = self.x_fine_end + 10 * self.dx * np.arange(1, self.n_coarse + 1)
        return np.concatenate([fine, coarse])

    def lookup(self, s: torch.Tensor, pair: int):
        """(f2 [n, 1], coefficient rows [n, E, K]) for normalized s values of one type pair."""
        sd = s.detach().to(torch.float64)
        kf = torch.floor((sd - self.x0) / self.dx)
        kc = self.n_fine + torch.floor((sd - self.x_fine_end) / (10 * self.dx))
        k = torch.where(sd < self.x_fine_end, kf, kc).clamp(0, self.n_fine + self.n_coarse - 1).to(torch.int64)
        knot = torch.where(k < self.n_fine, self.x0 + k.to(torch.float64) * self.dx,
                           self.x_fine_end + (k - self.n_fine).to(torch.float64) * (10 * self.dx))
        f2 = (s - knot.to(s.dtype)).reshape(-1, 1)
        return f2, self.table[pair][k.reshape(-1)]


def _hermite(y0, y1, d0, d1, h, s0=None, s1=None):
    """Coefficients (highest power first) of the polynomial on [0, h] matching values, first
    (and, for quintic, second) derivatives at both ends (cubic / quintic Hermite interpolation)."""
    D = y1 - y0
    if s0 is None:
        a = ((d0 + d1) * h - 2 * D) / h ** 3
        b = (3 * D - (2 * d0 + d1) * h) / h ** 2
        return [a, b, d0, y0]
    a = (12 * D - 6 * (d1 + d0) * h + (s1 - s0) * h ** 2) / (2 * h ** 5)
    b = (-30 * D + (14 * d1 + 16 * d0) * h + (3 * s0 - 2 * s1) * h ** 2) / (2 * h ** 4)
    c = (20 * D - (8 * d1 + 12 * d0) * h - (3 * s0 - s1) * h ** 2) / (2 * h ** 3)
    return [a, b, c, s0 / 2, d0, y0]


def build_table(emb_nets, davg, dstd, s_max_raw: float, dx: float = 0.01, order: int = 3,
                dtype=torch.float64, device="cpu") -> CompressTable:
    """Tabulates ``emb_nets`` (T*T callables s [n, 1] -> G [n, E], differentiable, float64) over the
    normalized s range of all centre types (set_sij_range, ckpt_compress.py:348-377)."""
    if order not in (3, 5):
        raise ValueError("compress order must be 3 or 5")
    davg = np.asarray(davg, np.float64)
    dstd = np.asarray(dstd, np.float64)
    s_max = max(0.0, float(np.max((s_max_raw - davg[:, 0]) / dstd[:, 0])))
    s_min = min(0.0, float(np.min((0.0 - davg[:, 0]) / dstd[:, 0])))
    n_fine = max(1, int(math.ceil((s_max - s_min) / dx)))
    x_end = s_min + n_fine * dx
    n_coarse = max(1, int(math.ceil((10 * s_max - x_end) / (10 * dx))))
    ct = CompressTable(None, s_min, dx, n_fine, n_coarse, order)
    x = torch.as_tensor(ct.knots(), dtype=torch.float64, device=device).unsqueeze(-1)
    h = torch.as_tensor(np.diff(ct.knots()), dtype=torch.float64, device=device).unsqueeze(-1)
    tabs = []
    for net in emb_nets:
        xs = x.clone().requires_grad_()
        y = net(xs)
        E = y.shape[-1]
        d1, d2 = [], []
        for i in range(E):
            (g,) = torch.autograd.grad(y[:, i].sum(), xs, retain_graph=True, create_graph=order == 5)
            d1.append(g[:, 0])
            if order == 5:
                (g2,) = torch.autograd.grad(g.sum(), xs, retain_graph=True)
                d2.append(g2[:, 0])
        y = y.detach()
        D1 = torch.stack(d1, -1).detach()
        D2 = torch.stack(d2, -1).detach() if order == 5 else None
        args = [y[:-1], y[1:], D1[:-1], D1[1:], h]
        if order == 5:
            args += [D2[:-1], D2[1:]]
        coef = torch.stack(_hermite(*args), -1)          # [rows, E, K]
        tabs.append(coef)
    ct.table = torch.stack(tabs).to(dtype)
    return ct
