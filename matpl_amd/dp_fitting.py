"""Host packing and launch of the matrix-core DP fitting kernel (csrc/dp_fitting.hip).

The fitting net of each centre type (src/model/dp_embedding.py:96-163) is
packed into the per-lane operand layouts the kernel consumes; atoms of each
type are split into tiles of 32. ``FittingMFMA(A)`` returns Ei and dEtot/dA.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .calcops import _ptr, _stream
from .dpnet import DPWeights

M2_FIXED = 16


def feat(r, h):
    return 8 * (r // 4) + 4 * h + (r % 4)


def _lane_maps():
    lanes = np.arange(64)
    h = lanes // 32
    i = lanes % 32
    v = np.arange(32)
    # input feature of value v in lane half h: (v // 16) * 32 + feat(v % 16, h)
    fin = (v[None, :] // 16) * 32 + 8 * ((v[None, :] % 16) // 4) + 4 * h[:, None] + (v[None, :] % 4)
    return h, i, fin


def _get(W, rows, cols):
    out = np.zeros(np.broadcast(rows, cols).shape, np.float64)
    ok = (rows < W.shape[0]) & (cols < W.shape[1])
    out[ok] = W[np.broadcast_to(rows, out.shape)[ok], np.broadcast_to(cols, out.shape)[ok]]
    return out


def segments(type_ranges):
    """[(type, first, end)] runs of atoms of one type: from [(first, end)] per type (index = type),
    or as given when already triples (frames packed along the atom axis, each sorted by type)."""
    rs = list(type_ranges)
    if rs and len(rs[0]) == 3:
        return [tuple(int(x) for x in r) for r in rs]
    return [(t, int(a), int(b)) for t, (a, b) in enumerate(rs)]


def type_runs(type_map: np.ndarray):
    """[(type, first, end)]: the maximal runs of equal type in a per-atom type array."""
    tm = np.asarray(type_map).reshape(-1)
    if tm.size == 0:
        return []
    cut = np.nonzero(np.diff(tm))[0] + 1
    starts = np.concatenate([[0], cut])
    ends = np.concatenate([cut, [tm.size]])
    return [(int(tm[a]), int(a), int(b)) for a, b in zip(starts, ends)]


class FittingMFMA:
    def __init__(self, w: DPWeights, type_ranges, precision="bf16", device="cuda"):
        cfg = w.cfg
        if cfg.M2 != M2_FIXED:
            raise ValueError("matrix-core fitting kernel needs M2 = 16")
        E = cfg.emb_out
        D = E * cfg.M2
        fit0 = w.fit[0]
        H = fit0[0][0].shape[1]
        L = len(fit0) - 1
        if H > 64 or any(l[0].shape[1] != H for l in fit0[1:-1]) or fit0[-1][0].shape[1] != 1 or L > 4:
            raise ValueError("matrix-core fitting kernel needs hidden widths <= 64 (equal) and <= 4 layers")
        self.E, self.D, self.H, self.L = E, D, H, L
        self.bf16 = precision == "bf16"
        Dh = D // 2
        nb0 = (D + 31) // 32
        hh, ii, fin = _lane_maps()
        T = cfg.ntypes
        w0f = np.zeros((T, 2, 64, Dh))
        w0b = np.zeros((T, nb0, 64, 32))
        whf = np.zeros((T, max(L - 1, 1), 2, 64, 32))
        whb = np.zeros((T, max(L - 1, 1), 2, 64, 32))
        nvec = L * 3 * 2 * 2 * 16 + 2 * 2 * 16 + 16
        vec = np.zeros((T, nvec))
        kk = np.arange(Dh)
        for t in range(T):
            layers = w.fit[t]
            W0, b0, _ = layers[0]
            for ob in range(2):
                out = 32 * ob + ii[:, None]
                if self.bf16:
                    k = 16 * (kk[None, :] // 8) + 8 * hh[:, None] + (kk[None, :] % 8)
                else:
                    k = 2 * kk[None, :] + hh[:, None]
                w0f[t, ob] = _get(W0, k, out)
            for blk in range(nb0):
                w0b[t, blk] = _get(W0, 32 * blk + ii[:, None], fin)
            for l in range(1, L):
                Wl = layers[l][0]
                for ob in range(2):
                    whf[t, l - 1, ob] = _get(Wl, fin, 32 * ob + ii[:, None])
                    whb[t, l - 1, ob] = _get(Wl, 32 * ob + ii[:, None], fin)
            v = vec[t, :L * 3 * 2 * 2 * 16].reshape(L, 3, 2, 2, 16)
            for l in range(L):
                b = layers[l][1]
                r = layers[l][2]
                for blk in range(2):
                    for h in range(2):
                        for rr in range(16):
                            f = 32 * blk + feat(rr, h)
                            if f < H:
                                v[l, 0, blk, h, rr] = b[f] if b is not None else 0.0
                                v[l, 1, blk, h, rr] = r[f] if r is not None else 1.0
            Wout, bout, _ = layers[L]
            o = vec[t, L * 3 * 2 * 2 * 16:L * 3 * 2 * 2 * 16 + 64].reshape(2, 2, 16)
            for blk in range(2):
                for h in range(2):
                    for rr in range(16):
                        f = 32 * blk + feat(rr, h)
                        if f < H:
                            o[blk, h, rr] = Wout[f, 0]
            vec[t, L * 3 * 2 * 2 * 16 + 64] = bout[0] if bout is not None else 0.0
        dt = torch.bfloat16 if self.bf16 else torch.float32
        self.w0f = torch.as_tensor(w0f.astype(np.float32), device=device).to(dt).contiguous()
        self.w0b = torch.as_tensor(w0b.astype(np.float32), device=device).to(dt).contiguous()
        self.whf = torch.as_tensor(whf.astype(np.float32), device=device).to(dt).contiguous()
        self.whb = torch.as_tensor(whb.astype(np.float32), device=device).to(dt).contiguous()
        self.vec = torch.as_tensor(vec.astype(np.float32), device=device).contiguous()
        self.strides = (w0f[0].size, w0b[0].size, whf[0].size, vec[0].size)
        tiles = []
        for t, a, b in segments(type_ranges):
            for s in range(a, b, 32):
                tiles.append((t, s, min(32, b - s)))
        self.ntiles = len(tiles)
        self.tiles = torch.as_tensor(np.asarray(tiles, np.int32).reshape(-1, 3), device=device).contiguous()

    def __call__(self, A: torch.Tensor):
        N = A.shape[0]
        Ei = torch.empty(N, dtype=torch.float32, device=A.device)
        dA = torch.empty_like(A)
        A = A.contiguous()
        _lib.call("matpl_dp_fitting", _ptr(A), _ptr(self.tiles), self.ntiles, _ptr(self.w0f), _ptr(self.w0b),
                  _ptr(self.whf), _ptr(self.whb), _ptr(self.vec), *self.strides, self.E, self.L, int(self.bf16),
                  _ptr(Ei), _ptr(dA), _stream())
        return Ei, dA
