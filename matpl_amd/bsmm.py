"""Batched small-matrix products on the HIP kernels of csrc/bsmm.hip, differentiable to any order.

``bmm_tn(X, Y)`` = X^T Y per batch element (X [n, K, a], Y [n, K, b] -> [n, a, b]) is the per-atom
contraction of the descriptor (A = Ri^T G over a bucket of slots, DR = A^T A[:, :, :M2]; the
reference's batched torch.matmul, src/model/dp_dp.py:332, 369, 412). The three products
  TN: X^T Y      NT: Y Z^T      NN: X Z
are closed under differentiation: the gradients of each are the other two (and itself), so the
autograd Functions below call one another in their backward passes and every derivative order of a
force-loss training step stays on these kernels.
"""
from __future__ import annotations

import torch

from . import _lib
from .calcops import _ptr, _stream


def _run(mode: int, P1, P2, shape):
    P1 = P1.contiguous()
    P2 = P2.to(P1.dtype).contiguous()
    if P1.dtype not in (torch.float32, torch.float64):
        raise RuntimeError(f"bsmm: unsupported dtype {P1.dtype}")
    out = torch.empty(shape, dtype=P1.dtype, device=P1.device)
    n = shape[0]
    if mode == 0:
        K, a, b = P1.shape[1], P1.shape[2], P2.shape[2]
    elif mode == 1:
        K, b, a = P1.shape[1], P1.shape[2], P2.shape[1]
    else:
        K, a, b = P1.shape[1], P1.shape[2], P2.shape[2]
    _lib.call("matpl_bsmm", int(P1.dtype == torch.float64), mode, _ptr(P1), _ptr(P2), _ptr(out), n, K, a, b,
              _stream())
    return out


def _check(X, Y, what):
    if X.dim() != 3 or Y.dim() != 3 or X.shape[0] != Y.shape[0]:
        raise ValueError(f"bsmm {what}: operands must be [n, ., .] with equal n, got {tuple(X.shape)} "
                         f"and {tuple(Y.shape)}")


class _TN(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, Y):
        _check(X, Y, "TN")
        if X.shape[1] != Y.shape[1]:
            raise ValueError("bsmm TN: X [n, K, a] and Y [n, K, b] need the same K")
        ctx.save_for_backward(X, Y)
        return _run(0, X, Y, (X.shape[0], X.shape[2], Y.shape[2]))

    @staticmethod
    def backward(ctx, g):
        X, Y = ctx.saved_tensors
        gX = _NT.apply(Y, g) if ctx.needs_input_grad[0] else None      # Y g^T  [n, K, a]
        gY = _NN.apply(X, g) if ctx.needs_input_grad[1] else None      # X g    [n, K, b]
        return gX, gY


class _NT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Y, Z):
        _check(Y, Z, "NT")
        if Y.shape[2] != Z.shape[2]:
            raise ValueError("bsmm NT: Y [n, K, b] and Z [n, a, b] need the same b")
        ctx.save_for_backward(Y, Z)
        return _run(1, Y, Z, (Y.shape[0], Y.shape[1], Z.shape[1]))

    @staticmethod
    def backward(ctx, g):
        Y, Z = ctx.saved_tensors
        gY = _NN.apply(g, Z) if ctx.needs_input_grad[0] else None      # g Z    [n, K, b]
        gZ = _TN.apply(g, Y) if ctx.needs_input_grad[1] else None      # g^T Y  [n, a, b]
        return gY, gZ


class _NN(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, Z):
        _check(X, Z, "NN")
        if X.shape[2] != Z.shape[1]:
            raise ValueError("bsmm NN: X [n, K, a] and Z [n, a, b] need the same a")
        ctx.save_for_backward(X, Z)
        return _run(2, X, Z, (X.shape[0], X.shape[1], Z.shape[2]))

    @staticmethod
    def backward(ctx, g):
        X, Z = ctx.saved_tensors
        gX = _NT.apply(g, Z) if ctx.needs_input_grad[0] else None      # g Z^T  [n, K, a]
        gZ = _TN.apply(X, g) if ctx.needs_input_grad[1] else None      # X^T g  [n, a, b]
        return gX, gZ


def bmm_tn(X, Y):
    """X^T Y per batch element: X [n, K, a], Y [n, K, b] -> [n, a, b]."""
    return _TN.apply(X, Y)


def bmm_nt(Y, Z):
    """Y Z^T per batch element: Y [n, K, b], Z [n, a, b] -> [n, K, a]."""
    return _NT.apply(Y, Z)


def bmm_nn(X, Z):
    """X Z per batch element: X [n, K, a], Z [n, a, b] -> [n, K, b]."""
    return _NN.apply(X, Z)
