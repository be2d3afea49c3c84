"""Products of a tall matrix (rows = atoms or neighbor slots) with a small weight matrix, for
nets trained through autograd.

hipBLAS picks a single-tile kernel for x^T g when x and g are tall and thin
(the weight gradient of a layer over all atoms / slots): one tile loops over
all K rows serially. ``_TallMM`` replaces x @ W so that the weight gradient,
at every derivative order, goes through ``_tall_tn``, which splits K into row
blocks of a batched product and sums them. In ``tall_mm_f32`` and
``tall_mm_bf16`` the products x @ W themselves, for FP32 device tensors with
a, b <= 64 (every DP embedding layer and the fitting net's hidden layers,
forward and input gradients), run on the library's tall-by-small kernel
(matpl_tall_mm), which reads and writes FP32 and, in BF16 mode, rounds operands
and result in registers; wider products stay with torch.
"""
from __future__ import annotations

import torch

from . import _lib
from .calcops import _ptr, _stream

_SPLIT_ROWS = 512
_SMALL_MAX = 64


def _small(x, W) -> bool:
    """x [K, a] @ W [a, b] goes to the HIP tall-by-small kernel: FP32 device tensors, a, b <= 64."""
    return (x.is_cuda and x.dtype == torch.float32 and W.dtype == torch.float32 and x.dim() == 2
            and 1 <= x.shape[1] <= _SMALL_MAX and 1 <= W.shape[1] <= _SMALL_MAX)


def _small_mm(x, W, lowp: bool):
    """x @ W on matpl_tall_mm (FP32 in and out; lowp: BF16 operands and result, FP32 sums). W may be
    any strided view (a transposed weight needs no copy); x needs unit column stride."""
    if x.stride(1) != 1 or x.stride(0) < x.shape[1]:
        x = x.contiguous()
    K, a = x.shape
    y = torch.empty((K, W.shape[1]), dtype=torch.float32, device=x.device)
    _lib.call("matpl_tall_mm", _ptr(x), max(x.stride(0), a), _ptr(W), W.stride(0), W.stride(1), _ptr(y), K, a,
              W.shape[1], int(lowp), _stream())
    return y


def _tall_tn(x, g):
    """x^T g for tall x [K, a] and g [K, b] with K >> a, b: the reduction over K is split into
    row blocks of a batched product and summed (one GEMM tile looping over all K is serial).
    Built from differentiable ops, so its own backward is two tall-by-small products."""
    K = x.shape[0]
    if K <= 2 * _SPLIT_ROWS:
        return x.t() @ g
    nblk = K // _SPLIT_ROWS
    Kf = nblk * _SPLIT_ROWS
    xs = x[:Kf].reshape(nblk, _SPLIT_ROWS, x.shape[1]).transpose(1, 2)
    gs = g[:Kf].reshape(nblk, _SPLIT_ROWS, g.shape[1])
    out = torch.bmm(xs, gs).sum(0)
    if Kf < K:
        out = out + x[Kf:].t() @ g[Kf:]
    return out


class _TallMM(torch.autograd.Function):
    """x @ W for tall x [K, a], small W [a, b]; the weight gradient goes through _tall_tn and the
    input gradient through _TallMM again, so every product of any derivative order is either
    tall-by-small or split over K. hip: the products x @ W that qualify run on matpl_tall_mm."""

    @staticmethod
    def forward(ctx, x, W, hip):
        ctx.save_for_backward(x, W)
        ctx.hip = hip
        return _small_mm(x, W, False) if hip and _small(x, W) else x @ W

    @staticmethod
    def backward(ctx, g):
        x, W = ctx.saved_tensors
        gx = _TallMM.apply(g, W.t(), ctx.hip) if ctx.needs_input_grad[0] else None
        gW = _tall_tn(x, g) if ctx.needs_input_grad[1] else None
        return gx, gW, None


def tall_mm(x, W):
    """x @ W for x [..., a] (any leading shape) and W [a, b]."""
    lead = x.shape[:-1]
    return _TallMM.apply(x.reshape(-1, x.shape[-1]), W, False).reshape(*lead, W.shape[-1])


def tall_mm_f32(x, W):
    """tall_mm whose FP32 products with a, b <= 64 run on matpl_tall_mm (for row counts in the
    hundreds of thousands, as in the batched DP training step; torch's product is quicker for a few
    thousand rows)."""
    lead = x.shape[:-1]
    return _TallMM.apply(x.reshape(-1, x.shape[-1]), W, True).reshape(*lead, W.shape[-1])


class _TallMMLow(torch.autograd.Function):
    """x @ W with both operands rounded to ``dt`` (bfloat16): BF16-rounded operands multiplied
    with FP32 VALU FMAs (matpl_tall_mm for widths <= 64, else torch's product), result rounded to
    ``dt`` and returned as FP32; the input gradient is the same low-precision product, the weight
    gradient an FP32 split product."""

    @staticmethod
    def forward(ctx, x, W, dt):
        ctx.save_for_backward(x, W)
        ctx.dt = dt
        if dt == torch.bfloat16 and _small(x, W):
            return _small_mm(x, W, True)
        return (x.to(dt) @ W.to(dt)).float()

    @staticmethod
    def backward(ctx, g):
        x, W = ctx.saved_tensors
        gx = _TallMMLow.apply(g, W.t(), ctx.dt) if ctx.needs_input_grad[0] else None
        gW = _tall_tn(x.float(), g.float()) if ctx.needs_input_grad[1] else None
        return gx, gW, None


def tall_mm_bf16(x, W):
    """x @ W, BF16 operands and product, FP32 accumulation; x [..., a], W [a, b]."""
    lead = x.shape[:-1]
    return _TallMMLow.apply(x.reshape(-1, x.shape[-1]), W, torch.bfloat16).reshape(*lead, W.shape[-1])



_SUM_ROWS = 256


def colsum(g):
    """Sum over all leading dimensions of g [..., b] -> [b], in two stages over row blocks (torch's
    single reduction over a tall dimension into a few outputs runs on a handful of threads)."""
    g2 = g.reshape(-1, g.shape[-1])
    K = g2.shape[0]
    if K <= 4 * _SUM_ROWS:
        return g2.sum(0)
    nblk = K // _SUM_ROWS
    Kf = nblk * _SUM_ROWS
    out = g2[:Kf].reshape(nblk, _SUM_ROWS, g2.shape[1]).sum(1).sum(0)
    if Kf < K:
        out = out + g2[Kf:].sum(0)
    return out


class _BiasAdd(torch.autograd.Function):
    """y + b (b broadcast over the rows of y); the bias gradient is colsum(g), itself built from
    differentiable ops."""

    @staticmethod
    def forward(ctx, y, b):
        return y + b

    @staticmethod
    def backward(ctx, g):
        return g, colsum(g)


def add_bias(y, b):
    return _BiasAdd.apply(y, b)
