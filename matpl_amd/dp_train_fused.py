"""Batched DP training step on hand-written kernels: forward, forces and the force-loss double
backward of the embedding nets on the matrix cores (csrc/dp_fused.hip, csrc/dp_train.hip).

The reference trains DP (src/PWMLFF/dp_mods/dp_trainer.py:163-275) by
dE = autograd.grad(Etot, Ri, create_graph=True) (src/model/dp_dp.py:429), forces from
calculateForce and loss.backward() through all of it (its force-loss edge is
calculateForceGrad.cu:9-90). Here, for a packed batch of frames (each with its own lattice, type
list and neighbor list with image shifts):

  1. matpl_dp_fused_forward: A_i = c_i (sum_j Ri_ij (x) G(s_ij) + empty slots), BF16 or FP32
     embedding on the matrix cores, c_i = 1 / (m * type_nums of i's frame) (dp_dp.py:413);
  2. the fitting nets in torch: E_i, and Abar = dE/dA with create_graph (dp_dp.py:304-342);
  3. matpl_dp_fused_backward + gather: forces F (atomic-free);
  4. the loss L = pe mean_f ((E_f - E^_f) / N_f)^2 + pf mean (F - F^)^2 and the force residual
     r = dL/dF;
  5. matpl_dp_train_tangent: dA = the tangent of A along r (forward mode through the nets);
  6. torch, through the fitting nets only: Ahat = d(L_E - <Abar, dA>)/dA and the fitting-net
     gradients (the force loss enters the fitting side as -<Abar, dA>);
  7. matpl_dp_train_embgrad: the embedding-net gradients from (Ahat, Abar, r), recomputing the
     nets' activations (nothing per neighbor is stored between the kernels);
  8. torch: the gradient of the empty-slot term sum (m - n_ij) Ri_pad (x) G(s_pad) (one net
     evaluation per type pair).
Loss and gradients equal the plain FP64 path's (DPTorch with create_graph, matpl_amd.dp_train.
plain_grads) within each precision mode's rounding; tests/test_gpu_dp_train_fused.py checks them,
and the plain path against central differences of the oracle's loss.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List

import numpy as np
import torch

from . import _lib
from .calcops import _ptr, _stream
from .dp import _Net
from .dp_fused import M_W2B, M_W2F, M_W3B, M_W3F, NMAT, NTI, NVEC, TANH_K, V_B1, V_B2, V_B3, V_GPAD, V_R2, V_R3, \
    V_W1, feat, type_masks
from .dpnet import DPWeights
from .neighbor import find_neighbore_batch, reverse_map


@dataclass
class FusedBatch:
    """Frames packed along the atom axis in the fused kernels' layout, with labels."""
    nframes: int
    natoms: int
    pos: torch.Tensor            # [N, 3] float64, wrapped into each frame's cell
    lattice: torch.Tensor        # [F, 9] float64
    type_map: torch.Tensor       # [N] int32
    atom_frame: torch.Tensor     # [N] int32
    frame_start: torch.Tensor    # [F + 1] int64
    frame_natoms: torch.Tensor   # [F] float64
    masks: torch.Tensor          # [F] int32 bit masks of each frame's type list
    type_present: torch.Tensor   # [F, T] bool
    by_type: List[torch.Tensor]  # atoms of each type (int64)
    nl_list: torch.Tensor        # [N, T*m] int32, 1-based within the frame
    nl_shift: torch.Tensor       # [N, T*m] int32 packed image shifts
    nl_counts: torch.Tensor      # [N, T] int32
    rev: torch.Tensor            # [N, T*m] int32 partner slot
    E_label: torch.Tensor        # [F] float64
    F_label: torch.Tensor        # [N, 3] float64


def pack_fused_batch(frames, cfg, device="cuda", seed=0, type_present=None) -> FusedBatch:
    """The neighbor lists (cutoff rcut, m per type, image shifts) of all frames in one launch, the
    reverse slot map, and seeded synthetic labels drawn as matpl_amd.dp_train.pack_batch draws them
    (same seed, same labels)."""
    T, m = cfg.ntypes, cfg.m
    nl = find_neighbore_batch([f.lattice for f in frames], [f.frac() for f in frames], [f.types for f in frames],
                              T, m, cfg.rcut, want_dR=False, want_shift=True, device=device)
    sizes = [f.natoms for f in frames]
    N = sum(sizes)
    fs = torch.as_tensor(np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64), device=device)
    af = torch.as_tensor(np.repeat(np.arange(len(frames), dtype=np.int32), sizes), device=device)
    lat = torch.as_tensor(np.stack([np.asarray(f.lattice, np.float64).reshape(9) for f in frames]), device=device)
    pos = torch.cat([nl.frac_wrapped[int(fs[k]):int(fs[k + 1])] @ lat[k].reshape(3, 3)
                     for k in range(len(frames))]).contiguous()
    tm = torch.as_tensor(np.concatenate([f.types for f in frames]).astype(np.int32), device=device)
    rev, rst = reverse_map(nl, tm, m, frame_start=fs, atom_frame=af)
    if int(rst.item()) != 0:
        raise RuntimeError("pack_fused_batch: neighbor list is not symmetric")
    if type_present is None:
        tp = np.zeros((len(frames), T), bool)
        for k, f in enumerate(frames):
            tp[k, np.unique(f.types)] = True
    else:
        tp = np.asarray(type_present, bool).reshape(len(frames), T)
        for k, f in enumerate(frames):
            if not tp[k, np.unique(f.types)].all():
                raise ValueError(f"type_present: frame {k} has atoms of a type outside its list")
    rng = np.random.default_rng(seed)
    E_label = torch.as_tensor(rng.normal(0.0, 0.1, len(frames)) * np.asarray(sizes), device=device)
    F_label = torch.as_tensor(rng.normal(0.0, 0.5, (N, 3)), device=device)
    tm64 = tm.to(torch.int64)
    return FusedBatch(len(frames), N, pos, lat, tm, af, fs, torch.as_tensor(sizes, dtype=torch.float64, device=device),
                      torch.as_tensor(type_masks(tp, len(frames), T).view(np.int32), device=device),
                      torch.as_tensor(tp, device=device), [torch.nonzero(tm64 == t).reshape(-1) for t in range(T)],
                      nl.list, nl.shift, nl.counts, rev, E_label, F_label)


class _EmbPack:
    """Gathers the embedding parameters (torch, on the device) into the fused kernels' per-pair
    operand layouts (matpl_amd.dp_fused.pack_fused's layout, built on the device every step).
    The parameters of all pairs are concatenated kind by kind (W1, b1, W2, b2, W3, b3, then G at
    the empty-slot input, a 0 and a 1), so the stacked per-kind arrays are views of one vector and
    each packed array is one gather from it through an index map fixed at construction, times a
    fixed factor (TANH_K on the forward weights and biases in BF16 mode)."""

    def __init__(self, E: int, P: int, K: float, device):
        self.E, self.P = E, P
        h = np.repeat([0, 1], 16)
        r = np.tile(np.arange(16), 2)
        f = np.array([feat(rr, hh) for hh, rr in zip(h, r)]).reshape(2, 16)
        mv = f < E                                                                  # [2, 16]
        fv = np.where(mv, f, 0)
        lanes = np.arange(64)
        hl, il = lanes // 32, lanes % 32
        fm = np.array([[feat(rr, hh) for rr in range(16)] for hh in hl])          # [64, 16]
        im = np.repeat(il[:, None], 16, 1)
        ok = (fm < E) & (im < E)
        fm, im = np.where(ok, fm, 0), np.where(ok, im, 0)
        PE, PEE = P * E, P * E * E
        o = {"W1": 0, "b1": PE, "W2": 2 * PE, "b2": 2 * PE + PEE, "W3": 3 * PE + PEE, "b3": 3 * PE + 2 * PEE,
             "g": 4 * PE + 2 * PEE}
        self.off = o
        ZERO, ONE = 5 * PE + 2 * PEE, 5 * PE + 2 * PEE + 1
        self.nflat = ONE + 1
        pp = np.arange(P)[:, None, None]
        iv = np.full((P, NVEC, 2, 16), ZERO, np.int64)
        kv = np.ones((P, NVEC, 2, 16), np.float32)
        for slot, key, scaled in ((V_W1, "W1", True), (V_B1, "b1", True), (V_B2, "b2", True), (V_B3, "b3", True),
                                  (V_GPAD, "g", False)):
            iv[:, slot] = np.where(mv[None], o[key] + pp * E + fv[None], ZERO)
            if scaled:
                kv[:, slot] = K
        for slot in (V_R2, V_R3):
            iv[:, slot] = np.where(mv[None], ONE, ZERO)
        im_ = np.full((P, NMAT, 64, 16), ZERO, np.int64)
        km = np.ones((P, NMAT, 64, 16), np.float32)
        for slot, key, fwd, scaled in ((M_W2F, "W2", True, True), (M_W3F, "W3", True, True),
                                       (M_W2B, "W2", False, False), (M_W3B, "W3", False, False)):
            row, col = (fm, im) if fwd else (im, fm)                   # fwd: W[f(r,h)][lane % 32]
            im_[:, slot] = np.where(ok[None], o[key] + pp * E * E + row[None] * E + col[None], ZERO)
            if scaled:
                km[:, slot] = K
        self.i_vec = torch.as_tensor(iv, device=device)
        self.k_vec = torch.as_tensor(kv, device=device)
        self.i_mat = torch.as_tensor(im_, device=device)
        self.k_mat = torch.as_tensor(km, device=device)
        self.const = torch.tensor([0.0, 1.0], dtype=torch.float32, device=device)

    def stacked(self, flat):
        """W1, b1, W2, b2, W3, b3 of all pairs ([P, E] / [P, E, E]) as views of the concatenation."""
        P, E, o = self.P, self.E, self.off
        v = lambda k, n: flat[o[k]:o[k] + n]
        return (v("W1", P * E).view(P, E), v("b1", P * E).view(P, E), v("W2", P * E * E).view(P, E, E),
                v("b2", P * E).view(P, E), v("W3", P * E * E).view(P, E, E), v("b3", P * E).view(P, E))

    def pack(self, flat, gpad):
        """(vec [P, NVEC, 2, 16], mat [P, NMAT, 64, 16]) in float32."""
        full = torch.cat([flat, gpad.reshape(-1), self.const])
        return full[self.i_vec] * self.k_vec, full[self.i_mat] * self.k_mat


class _FitPack:
    """The fitting nets' parameters (torch, on the device) in the matrix-core fitting kernels' operand
    layouts (matpl_amd.dp_fitting.FittingMFMA's packing), built on the device every step: the
    parameters are concatenated into one vector (plus a 0 and a 1) and each packed array is one
    gather from it through an index map fixed at construction."""

    def __init__(self, E: int, H: int, L: int, T: int, has_rdt, bf16: bool, device):
        from .dp_fitting import _lane_maps
        self.E, self.H, self.L, self.T, self.bf16 = E, H, L, T, bf16
        D = 16 * E
        self.D, self.Dh, self.nb0 = D, D // 2, (D + 31) // 32
        hh, ii, fin = _lane_maps()
        kk = np.arange(self.Dh)
        # offsets of each type's parameters in the concatenation, in the order of _flat_params
        self.has_rdt = list(has_rdt)
        sizes = [D * H, H] + [H * H] * (L - 1) + [H] * (L - 1) + [H * int(r) for r in self.has_rdt[1:L]] + [H, 1]
        per = int(sum(sizes))
        base = np.arange(T)[:, None, None, None] * per
        off = np.cumsum([0] + sizes)
        o_W0, o_b0 = off[0], off[1]
        o_Wh = [off[2 + (l - 1)] for l in range(1, L)]
        o_bh = [off[2 + (L - 1) + (l - 1)] for l in range(1, L)]
        o_rh = [off[2 + 2 * (L - 1) + (l - 1)] for l in range(1, L)]
        o_wo, o_bo = off[-3], off[-2]
        ZERO, ONE = T * per, T * per + 1
        self.nflat = T * per + 2

        def pick(valid, idx):
            return np.where(valid, idx, ZERO)
        k = (16 * (kk[None, :] // 8) + 8 * hh[:, None] + (kk[None, :] % 8)) if bf16 else (2 * kk[None, :] + hh[:, None])
        out = np.stack([32 * ob + ii[:, None] for ob in range(2)])                 # [2, 64, 1]
        w0f = pick((k[None] < D) & (out < H), base + o_W0 + k[None] * H + out)
        rows = np.stack([32 * blk + ii[:, None] for blk in range(self.nb0)])       # [nb0, 64, 1]
        w0b = pick((rows < D) & (fin[None] < H), base + o_W0 + rows * H + fin[None])
        self.i_w0f = torch.as_tensor(w0f.reshape(T, 2, 64, self.Dh), device=device)
        self.i_w0b = torch.as_tensor(w0b.reshape(T, self.nb0, 64, 32), device=device)
        nh = max(L - 1, 1)
        whf = np.full((T, nh, 2, 64, 32), ZERO, np.int64)
        whb = np.full((T, nh, 2, 64, 32), ZERO, np.int64)
        tb = np.arange(T)[:, None, None, None] * per
        for l in range(1, L):
            o = o_Wh[l - 1]
            whf[:, l - 1] = pick((fin[None] < H) & (out < H), tb + o + fin[None] * H + out)
            whb[:, l - 1] = pick((out < H) & (fin[None] < H), tb + o + out * H + fin[None])
        self.i_whf = torch.as_tensor(whf, device=device)
        self.i_whb = torch.as_tensor(whb, device=device)
        nvec = L * 3 * 2 * 2 * 16 + 2 * 2 * 16 + 16
        vec = np.full((T, nvec), ZERO, np.int64)
        f = np.array([[[32 * blk + feat(rr, h) for rr in range(16)] for h in range(2)] for blk in range(2)])   # [2,2,16]
        ok = f < H
        tt = np.arange(T)[:, None, None, None] * per
        v = vec[:, :L * 3 * 64].reshape(T, L, 3, 2, 2, 16)
        for l in range(L):
            ob_ = o_b0 if l == 0 else o_bh[l - 1]
            v[:, l, 0] = np.where(ok[None], tt + ob_ + f[None], ZERO)
            if l >= 1 and self.has_rdt[l]:
                v[:, l, 1] = np.where(ok[None], tt + o_rh[l - 1] + f[None], ZERO)
            else:
                v[:, l, 1] = np.broadcast_to(np.where(ok, ONE, ZERO), (T, 2, 2, 16))
        vec[:, L * 3 * 64:L * 3 * 64 + 64] = np.where(ok[None], tt + o_wo + f[None], ZERO).reshape(T, 64)
        vec[:, L * 3 * 64 + 64] = np.arange(T) * per + o_bo
        self.i_vec = torch.as_tensor(vec, device=device)
        self.const = torch.tensor([0.0, 1.0], dtype=torch.float32, device=device)
        self.strides = (2 * 64 * self.Dh, self.nb0 * 64 * 32, nh * 2 * 64 * 32, nvec)

    def _flat_params(self, nets):
        ps = []
        L = self.L
        for n in nets:
            ps += [n.W[0].reshape(-1), n.b[0]]
            ps += [n.W[l].reshape(-1) for l in range(1, L)]
            ps += [n.b[l] for l in range(1, L)]
            ps += [n.r[l] for l in range(1, L) if self.has_rdt[l]]
            ps += [n.W[L].reshape(-1), n.b[L].reshape(-1)]
        return ps

    def __call__(self, nets):
        """(w0f, w0b, whf, whb, vec, strides) for the list of per-type _Net fitting nets."""
        with torch.no_grad():
            flat = torch.cat(self._flat_params(nets) + [self.const])
            dt = torch.bfloat16 if self.bf16 else torch.float32
            w0f, w0b, whf, whb = (flat[i].to(dt) for i in (self.i_w0f, self.i_w0b, self.i_whf, self.i_whb))
            vec = flat[self.i_vec]
        return w0f, w0b, whf, whb, vec, self.strides


class DPTrainFused(torch.nn.Module):
    """The DP model's parameters (FP32, DPTorch's module layout) and its training step on the
    hand-written kernels. precision: "bf16" (BF16 matrix-core products, FP32 accumulation) or
    "fp32"."""

    def __init__(self, w: DPWeights, precision: str = "bf16", device="cuda", pe: float = 1.0, pf: float = 1.0,
                 fitting: str = "kernel"):
        """fitting: "kernel" (the fitting nets' energies, first and second derivatives on the
        matrix-core kernels matpl_dp_fitting_perm / matpl_dp_fit_train) or "torch" (torch autograd
        with create_graph)."""
        super().__init__()
        if fitting not in ("kernel", "torch"):
            raise ValueError("fitting must be 'kernel' or 'torch'")
        cfg = w.cfg
        if precision not in ("bf16", "fp32"):
            raise ValueError("precision must be 'bf16' or 'fp32'")
        es = cfg.emb_sizes
        if len(es) != 3 or es[1] != es[0] or es[2] != es[0] or es[0] not in (16, 25):
            raise ValueError("the training kernels need embedding sizes [E, E, E] with E in {16, 25}")
        if cfg.emb_resnet_dt:
            raise ValueError("the training kernels need an embedding without resnet_dt (the reference default)")
        self.cfg = cfg
        self.E = es[0]
        self.bf16 = precision == "bf16"
        self.pe, self.pf = pe, pf
        T = cfg.ntypes
        f32 = torch.float32
        self.emb = torch.nn.ModuleList([_Net(w.emb[p], 0, f32, device) for p in range(T * T)])
        self.fit = torch.nn.ModuleList([_Net(w.fit[t], 1, f32, device) for t in range(T)])
        dev = torch.device(device)
        tinfo = np.zeros((T, NTI), np.float32)
        ri_pad = -w.davg / w.dstd
        for ti in range(T):
            tinfo[ti, 0:4] = w.davg[ti]
            tinfo[ti, 4:8] = 1.0 / w.dstd[ti]
            tinfo[ti, 8:12] = w.dstd[ti]
            tinfo[ti, 12:16] = ri_pad[ti]
        self.register_buffer("tinfo", torch.as_tensor(tinfo, device=dev))
        self.register_buffer("ri_pad", torch.as_tensor(ri_pad, dtype=f32, device=dev))
        self._pk = _EmbPack(self.E, T * T, TANH_K if self.bf16 else 1.0, dev)
        self.gsize = _lib.query("matpl_dp_train_grad_size", self.E)
        self.chunk = 0          # atoms per wavefront of the gradient kernel (0: chosen from the batch)
        fit0 = w.fit[0]
        H, L = fit0[0][0].shape[1], len(fit0) - 1
        ok = (cfg.M2 == 16 and H <= 64 and 1 <= L <= 4 and all(l[0].shape[1] == H for l in fit0[1:-1]) and
              fit0[-1][0].shape[1] == 1 and all(b is not None for _, b, _ in fit0))
        if fitting == "kernel" and not ok:
            raise ValueError("the fitting kernels need M2 = 16, 1..4 hidden layers of one width <= 64, biases")
        self.fitting = fitting
        self.H, self.L = H, L
        self._fp = _FitPack(self.E, H, L, T, self.fit[0].has_rdt, self.bf16, dev) if fitting == "kernel" else None
        self._tiles = {}

    # ---------------------------------------------------------------- weights
    def _flat(self):
        """The embedding parameters of all type pairs in one vector, kind by kind (_EmbPack)."""
        ns = self.emb
        return torch.cat([n.W[0].reshape(-1) for n in ns] + [n.b[0] for n in ns] +
                         [n.W[1].reshape(-1) for n in ns] + [n.b[1] for n in ns] +
                         [n.W[2].reshape(-1) for n in ns] + [n.b[2] for n in ns])

    def _stacked(self):
        """The embedding parameters of all type pairs stacked: W1, b1, b2, b3 [P, E], W2, W3 [P, E, E]
        (returned as W1, b1, W2, b2, W3, b3; views of one concatenation)."""
        return self._pk.stacked(self._flat())

    def _gpad(self, W1, b1, W2, b2, W3, b3):
        """G(s_pad) of every pair [P, E] (the embedding layer rules: plain first layer, residual
        E -> E layers), batched over pairs."""
        T = self.cfg.ntypes
        x0 = self.ri_pad[torch.arange(T * T, device=W1.device) // T, 0]
        h = torch.tanh(x0[:, None] * W1 + b1)
        h = h + torch.tanh(torch.bmm(h[:, None, :], W2)[:, 0] + b2)
        return h + torch.tanh(torch.bmm(h[:, None, :], W3)[:, 0] + b3)

    def _pack(self):
        """(vec, mat) from the current parameters, and G at each pair's empty-slot input."""
        pk = self._pk
        with torch.no_grad():
            flat = self._flat()
            vec, mat = pk.pack(flat, self._gpad(*pk.stacked(flat)))
        return vec, (mat.to(torch.bfloat16) if self.bf16 else mat)

    # ---------------------------------------------------------------- fitting
    def _energies(self, A: torch.Tensor, b: FusedBatch):
        M2 = self.cfg.M2
        Ei = torch.zeros(b.natoms, dtype=A.dtype, device=A.device)
        for t, idx in enumerate(b.by_type):
            if idx.numel() == 0:
                continue
            At = A[idx]
            DR = (At.transpose(1, 2) @ At[:, :, :M2]).reshape(idx.numel(), -1)
            Ei = Ei.index_copy(0, idx, self.fit[t](DR).reshape(-1))
        return Ei

    # ---------------------------------------------------------------- one step
    def _head(self, b: FusedBatch):
        cfg = self.cfg
        m = cfg.m
        return (_ptr(b.pos), _ptr(b.lattice), _ptr(b.atom_frame), _ptr(b.frame_start), _ptr(b.masks),
                _ptr(b.nl_list), _ptr(b.nl_shift), _ptr(b.nl_counts), _ptr(b.type_map))

    def _tail(self, b: FusedBatch, mnl: int):
        cfg = self.cfg
        return (_ptr(self.tinfo), b.natoms, cfg.ntypes, mnl, cfg.m, self.E, int(self.bf16), float(cfg.rcut),
                float(cfg.rcut_smooth), 1.0 / (cfg.m * cfg.ntypes))

    def _listed_any(self, b: FusedBatch):
        """Per neighbor type: whether any frame of the batch lists it (cached per batch, host side)."""
        hit = getattr(self, "_listed_cache", None)
        if hit is not None and hit[0] is b:
            return hit[1]
        val = b.type_present.any(0).cpu().tolist()
        self._listed_cache = (b, val)
        return val

    def _fit_tiles(self, b: FusedBatch):
        """perm (atoms grouped by type, int32), tiles of <= 32 positions of one type, and each type's
        position range, for the fitting kernels (cached per batch)."""
        key = id(b)
        hit = self._tiles.get(key)
        if hit is not None and hit[0] is b:
            return hit[1]
        rows, ranges, off = [], [], 0
        for t, idx in enumerate(b.by_type):
            c = int(idx.numel())
            for s0 in range(0, c, 32):
                rows.append((t, off + s0, min(32, c - s0)))
            ranges.append((off, off + c))
            off += c
        perm = torch.cat(list(b.by_type)).to(torch.int32).contiguous()
        tiles = torch.as_tensor(np.asarray(rows, np.int32).reshape(-1, 3), device=perm.device)
        val = (perm, tiles, ranges)
        self._tiles = {key: (b, val)}
        return val

    def _fit_kernel_grads(self, b: FusedBatch, A, dA, gE, fpk):
        """Ahat = dL/dA and the fitting-net gradients (self.fit.parameters() order) from
        matpl_dp_fit_train's per-atom rows: weight gradients are sums of products over each type's
        atoms, bias / rdt / output gradients sums of rows."""
        w0f, w0b, whf, whb, vec, strides = fpk
        perm, tiles, ranges = self._fit_tiles(b)
        N, E, H, L, D = b.natoms, self.E, self.H, self.L, 16 * self.E
        dev = A.device
        f32 = torch.float32
        ahat = torch.empty_like(A)
        xw0 = torch.empty((N, 2, D), dtype=f32, device=dev)
        zw0 = torch.empty((N, 2, 64), dtype=f32, device=dev)
        nh = max(L - 1, 1)
        xh = torch.empty((N, nh, 2, 64), dtype=f32, device=dev)
        zh = torch.empty((N, nh, 2, 64), dtype=f32, device=dev)
        rb = torch.empty((N, nh, 64), dtype=f32, device=dev)
        xo = torch.empty((N, 64), dtype=f32, device=dev)
        tz = torch.empty((N, L, 2, 2, 32), dtype=f32, device=dev)
        _lib.call("matpl_dp_fit_train", _ptr(A), _ptr(dA), _ptr(gE), _ptr(perm), _ptr(tiles), tiles.shape[0],
                  _ptr(w0f), _ptr(w0b), _ptr(whf), _ptr(whb), _ptr(vec), *strides, E, L, int(self.bf16), _ptr(ahat),
                  _ptr(xw0), _ptr(zw0), _ptr(xh), _ptr(zh), _ptr(rb), _ptr(xo), _ptr(tz), _stream())
        gE_pos = gE[self._batch_const(b)["perm"]]
        # weight gradients: per type the products over its atoms' (value, tangent) rows, all types
        # in one launch per layer (matpl_tall_xtz)
        T = len(self.fit)
        start = np.asarray([r[0] for r in ranges] + [ranges[-1][1]], np.int64)
        maxa = int(max(e - a for a, e in ranges))
        # outputs [T][in][H]: each type's block is the contiguous gradient of its weight matrix (the
        # optimizer's multi-tensor kernels need gradients laid out like their parameters)
        work = torch.empty(max(_lib.query("matpl_tall_xtz_work", maxa, T, max(D, H), H), 1), dtype=f32, device=dev)
        Y0 = torch.empty((T, D, H), dtype=f32, device=dev)
        _lib.call("matpl_tall_xtz", _ptr(xw0), 2 * D, D, _ptr(zw0), 128, 64, start.ctypes.data, T, D, H, _ptr(Y0),
                  _ptr(work), work.numel(), _stream())
        Yh = []
        for l in range(1, L):
            Y = torch.empty((T, H, H), dtype=f32, device=dev)
            xp = xh.data_ptr() + (l - 1) * 128 * 4
            zp = zh.data_ptr() + (l - 1) * 128 * 4
            _lib.call("matpl_tall_xtz", xp, (L - 1) * 128, 64, zp, (L - 1) * 128, 64, start.ctypes.data, T, H, H,
                      _ptr(Y), _ptr(work), work.numel(), _stream())
            Yh.append(Y)
        # sums of the bias, resnet_dt and output rows over each type's positions, one launch:
        # zw0 value rows, zh value rows per layer, rb per layer, xo, gE
        segs = [(zw0.data_ptr(), 128, H)]
        segs += [(zh.data_ptr() + (l - 1) * 128 * 4, (L - 1) * 128, H) for l in range(1, L)]
        segs += [(rb.data_ptr() + (l - 1) * 64 * 4, (L - 1) * 64, H) for l in range(1, L)]
        segs += [(xo.data_ptr(), 64, H), (gE_pos.data_ptr(), 1, 1)]
        sptr = np.asarray([x[0] for x in segs], np.uint64)
        sld = np.asarray([x[1] for x in segs], np.int64)
        scol = np.asarray([x[2] for x in segs], np.int32)
        totc = int(scol.sum())
        swork = torch.empty(max(_lib.query("matpl_seg_rowsum_work", maxa, T, totc), 1), dtype=f32, device=dev)
        sums = torch.empty((T, totc), dtype=f32, device=dev)
        _lib.call("matpl_seg_rowsum", len(segs), sptr.ctypes.data, sld.ctypes.data, scol.ctypes.data,
                  start.ctypes.data, T, _ptr(sums), _ptr(swork), swork.numel(), _stream())
        cs = np.concatenate([[0], np.cumsum(scol)])
        col = lambda k: sums[:, int(cs[k]):int(cs[k + 1])]
        out = []
        for t, net in enumerate(self.fit):
            a, e = ranges[t]
            if e == a:
                out += [None] * (3 * len(net.W))
                continue
            gW = [Y0[t]]
            gb = [col(0)[t]]
            gr = [None]
            for l in range(1, L):
                gW.append(Yh[l - 1][t])
                gb.append(col(l)[t])
                gr.append(col(L - 1 + l)[t] if net.has_rdt[l] else None)
            gW.append(col(2 * L - 1)[t].reshape(H, 1))
            gb.append(col(2 * L)[t].reshape(1))
            gr.append(None)
            out += gW + gb + gr
        return ahat, out

    def loss_and_grads(self, b: FusedBatch):
        """(loss, gradients aligned with self.parameters(); None for parameters the loss does not reach)."""
        if self.fitting == "kernel":
            return self._loss_and_grads_kernels(b)
        cfg = self.cfg
        T, m, N, E = cfg.ntypes, cfg.m, b.natoms, self.E
        dev = b.pos.device
        mnl = b.nl_list.shape[1] // T
        vec, mat = self._pack()
        head, tail = self._head(b), self._tail(b, mnl)
        status = torch.zeros(1, dtype=torch.int32, device=dev)
        # 1. descriptor A
        A = torch.empty((N, 4, E), dtype=torch.float32, device=dev)
        _lib.call("matpl_dp_fused_forward", *head, _ptr(vec), _ptr(mat), *tail, _ptr(A), _ptr(status), _stream())
        # 2. fitting, Abar = dE/dA kept differentiable
        A.requires_grad_()
        Ei = self._energies(A, b)
        Ef = torch.zeros(b.nframes, dtype=Ei.dtype, device=dev).index_add(0, b.atom_frame.to(torch.int64), Ei)
        (Abar,) = torch.autograd.grad(Ei.sum(), A, create_graph=True)
        abar = Abar.detach().contiguous()
        # 3. forces
        F = torch.empty((N, 3), dtype=torch.float32, device=dev)
        V = torch.zeros((b.nframes, 9), dtype=torch.float64, device=dev)
        fslot = torch.empty((N * T * mnl, 4), dtype=torch.float32, device=dev)
        fcenter = torch.empty((N, 3), dtype=torch.float32, device=dev)
        _lib.call("matpl_dp_fused_backward", *head, _ptr(vec), _ptr(mat), *tail, _ptr(abar), _ptr(b.rev), _ptr(fslot),
                  _ptr(fcenter), _ptr(F), _ptr(V), _ptr(status), _stream())
        _lib.call("matpl_dp_fused_gather", _ptr(b.nl_counts), _ptr(fslot), _ptr(fcenter), _ptr(F), T, mnl, N, _stream())
        # 4. loss and the force residual r = dL/dF
        le = (((Ef.double() - b.E_label) / b.frame_natoms) ** 2).mean()
        lf = ((F.double() - b.F_label) ** 2).mean()
        resid = (2.0 * self.pf / (3 * N) * (F.double() - b.F_label)).to(torch.float32).contiguous()
        # 5. tangent of A along r
        dA = torch.empty((N, 4, E), dtype=torch.float32, device=dev)
        nin = torch.empty((N, T), dtype=torch.int32, device=dev)
        _lib.call("matpl_dp_train_tangent", *head, _ptr(vec), _ptr(mat), *tail, _ptr(resid), _ptr(dA), _ptr(nin),
                  _ptr(status), _stream())
        # 6. fitting side: the force loss enters as -<Abar, dA>
        fparams = [q for q in self.fit.parameters()]
        ltot = self.pe * le - (Abar * dA).sum()
        g = torch.autograd.grad(ltot, fparams + [A], allow_unused=True)
        gfit, ahat = list(g[:-1]), g[-1].detach().contiguous()
        # 7. embedding gradients
        grad = torch.zeros((T * T, self.gsize), dtype=torch.float32, device=dev)
        chunk = self.chunk or max(8, -(-N * T // 8192))
        _lib.call("matpl_dp_train_embgrad", *head, _ptr(vec), _ptr(mat), *tail, _ptr(resid), _ptr(abar), _ptr(ahat),
                  _ptr(grad), chunk, _ptr(status), _stream())
        gemb = self._emb_grads(grad, b, ahat, nin)
        loss = (self.pe * le + self.pf * lf).detach()
        # the overflow flag depends only on the batch's positions and lists, which training does not
        # change: it is read (a host synchronization) on the first step over a batch only, once
        # everything is enqueued, and later steps over the same batch run without a host stall
        if not getattr(b, "_lists_checked", False):
            if int(status.item()) != 0:
                raise OverflowError("more than maxNeighborNum neighbors of one type within rcut")
            b._lists_checked = True
        return loss, gemb + gfit

    def _loss_and_grads_kernels(self, b: FusedBatch):
        """The step with the fitting nets on the matrix-core kernels as well: energies and Abar from
        matpl_dp_fitting_perm, the fitting side's second derivatives from matpl_dp_fit_train."""
        cfg = self.cfg
        T, N, E = cfg.ntypes, b.natoms, self.E
        dev = b.pos.device
        mnl = b.nl_list.shape[1] // T
        vec, mat = self._pack()
        fpk = self._fp(list(self.fit))
        w0f, w0b, whf, whb, fvec, strides = fpk
        perm, tiles, _ = self._fit_tiles(b)
        head, tail = self._head(b), self._tail(b, mnl)
        status = torch.zeros(1, dtype=torch.int32, device=dev)
        # 1. descriptor A
        A = torch.empty((N, 4, E), dtype=torch.float32, device=dev)
        _lib.call("matpl_dp_fused_forward", *head, _ptr(vec), _ptr(mat), *tail, _ptr(A), _ptr(status), _stream())
        # 2. fitting: Ei and Abar = dEi/dA
        Ei = torch.empty(N, dtype=torch.float32, device=dev)
        abar = torch.empty_like(A)
        _lib.call("matpl_dp_fitting_perm", _ptr(A), _ptr(tiles), _ptr(perm), tiles.shape[0], _ptr(w0f), _ptr(w0b),
                  _ptr(whf), _ptr(whb), _ptr(fvec), *strides, E, self.L, int(self.bf16), _ptr(Ei), _ptr(abar),
                  _stream())
        k = self._batch_const(b)
        af = k["af"]
        Ef = torch.zeros(b.nframes, dtype=torch.float64, device=dev).index_add(0, af, Ei.double())
        # 3. forces
        F = torch.empty((N, 3), dtype=torch.float32, device=dev)
        V = torch.zeros((b.nframes, 9), dtype=torch.float64, device=dev)
        fslot = torch.empty((N * T * mnl, 4), dtype=torch.float32, device=dev)
        fcenter = torch.empty((N, 3), dtype=torch.float32, device=dev)
        _lib.call("matpl_dp_fused_backward", *head, _ptr(vec), _ptr(mat), *tail, _ptr(abar), _ptr(b.rev), _ptr(fslot),
                  _ptr(fcenter), _ptr(F), _ptr(V), _ptr(status), _stream())
        _lib.call("matpl_dp_fused_gather", _ptr(b.nl_counts), _ptr(fslot), _ptr(fcenter), _ptr(F), T, mnl, N, _stream())
        # 4. loss, dL/dEi of the energy term, and the force residual r = dL/dF
        de = (Ef - b.E_label) / b.frame_natoms
        le = (de ** 2).mean()
        lf = ((F.double() - b.F_label) ** 2).mean()
        gE = (self.pe * 2.0 / b.nframes * de / b.frame_natoms)[af].to(torch.float32).contiguous()
        resid = (2.0 * self.pf / (3 * N) * (F.double() - b.F_label)).to(torch.float32).contiguous()
        # 5. tangent of A along r
        dA = torch.empty((N, 4, E), dtype=torch.float32, device=dev)
        nin = torch.empty((N, T), dtype=torch.int32, device=dev)
        _lib.call("matpl_dp_train_tangent", *head, _ptr(vec), _ptr(mat), *tail, _ptr(resid), _ptr(dA), _ptr(nin),
                  _ptr(status), _stream())
        # 6. fitting side: Ahat = d(pe L_E - <Abar, dA>)/dA and the fitting-net gradients
        ahat, gfit = self._fit_kernel_grads(b, A, dA, gE, fpk)
        # 7. embedding gradients
        grad = torch.zeros((T * T, self.gsize), dtype=torch.float32, device=dev)
        chunk = self.chunk or max(8, -(-N * T // 8192))
        _lib.call("matpl_dp_train_embgrad", *head, _ptr(vec), _ptr(mat), *tail, _ptr(resid), _ptr(abar), _ptr(ahat),
                  _ptr(grad), chunk, _ptr(status), _stream())
        gemb = self._emb_grads(grad, b, ahat, nin)
        loss = (self.pe * le + self.pf * lf).detach()
        # the overflow flag depends only on the batch's positions and lists, which training does not
        # change: it is read (a host synchronization) on the first step over a batch only, once
        # everything is enqueued, and later steps over the same batch run without a host stall
        if not getattr(b, "_lists_checked", False):
            if int(status.item()) != 0:
                raise OverflowError("more than maxNeighborNum neighbors of one type within rcut")
            b._lists_checked = True
        return loss, gemb + gfit

    def _batch_const(self, b: FusedBatch):
        """Per-batch arrays the step reuses (a training batch's atoms and lists do not change): the
        frame index as int64, each atom's listed types and 1 / (m * their number), Ri_pad of each
        atom's type and the type-grouped atom order; kept on the batch for this model."""
        got = getattr(b, "_step_const", None)
        key = (self, self.ri_pad._version)         # an in-place change of Ri_pad (load_state_dict) renews it
        if got is not None and got[0][0] is self and got[0][1] == key[1]:
            return got[1]
        m = self.cfg.m
        af = b.atom_frame.to(torch.int64)
        listed = b.type_present[af].to(torch.float32)
        k = {"af": af, "listed": listed, "c": 1.0 / (m * listed.sum(1)),
             "rp": self.ri_pad[b.type_map.to(torch.int64)][:, :, None],
             "perm": self._fit_tiles(b)[0].to(torch.int64)}
        b._step_const = (key, k)
        return k

    def _emb_grads(self, grad, b: FusedBatch, ahat, nin):
        """Per-parameter embedding gradients: the kernel's sums plus the empty-slot term."""
        cfg, T, E, m = self.cfg, self.cfg.ntypes, self.E, self.cfg.m
        # empty slots: u[p] = sum_i c_i listed(i, tj) (m - n_ij) ahat_i^T Ri_pad(ti)
        k = self._batch_const(b)
        listed, c = k["listed"], k["c"]                                                # [N, T], [N]
        npad = (m - nin.to(torch.float32)) * listed                                     # [N, T]
        proj = (ahat * k["rp"]).sum(1)                                                  # [N, E]
        wsum = (c[:, None] * npad)[:, :, None] * proj[:, None, :]                        # [N, T, E]
        # per centre type sums over its atoms (one gather into type order, then contiguous sums; an
        # index_add onto T rows would serialize on atomics)
        _, _, ranges = self._fit_tiles(b)
        wp = wsum[k["perm"]]
        u = torch.stack([wp[a:e].sum(0) if e > a else torch.zeros((T, E), dtype=wsum.dtype, device=wsum.device)
                         for a, e in ranges])
        has_atoms = [bool(idx.numel()) for idx in b.by_type]
        listed_any = self._listed_any(b)
        # the empty-slot term's gradient, all pairs at once
        with torch.enable_grad():
            leaves = [x.detach().requires_grad_() for x in self._stacked()]
            gpad = self._gpad(*leaves)
            pg = torch.autograd.grad((u.reshape(T * T, E) * gpad).sum(), leaves)
        o = self._offsets()
        P = T * T
        gW1 = grad[:, o["W1"]:o["W1"] + E] + pg[0]
        gb1 = grad[:, o["B1"]:o["B1"] + E] + pg[1]
        gW2 = grad[:, o["W2"]:o["W2"] + E * E].reshape(P, E, E) + pg[2]
        gb2 = grad[:, o["B2"]:o["B2"] + E] + pg[3]
        gW3 = grad[:, o["W3"]:o["W3"] + E * E].reshape(P, E, E) + pg[4]
        gb3 = grad[:, o["B3"]:o["B3"] + E] + pg[5]
        out = []
        for p in range(P):
            a, t = divmod(p, T)
            net = self.emb[p]
            if not (has_atoms[a] and listed_any[t]):
                out += [None] * (3 * len(net.W))
                continue
            out += [gW1[p].view(1, E), gW2[p], gW3[p], gb1[p], gb2[p], gb3[p]]
            out += [None] * len(net.r)
        return out

    def _offsets(self):
        E = self.E
        W3, B3 = 0, E * E
        W2 = B3 + E
        B2 = W2 + E * E
        W1 = B2 + E
        return {"W3": W3, "B3": B3, "W2": W2, "B2": B2, "W1": W1, "B1": W1 + E}
