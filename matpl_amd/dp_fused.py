"""Fused DP evaluation: energy, forces and virial from positions + neighbor list.

Three stages per evaluation, all on the current stream:
  1. ``matpl_dp_fused_forward``: geometry (FP64 minimum image, FP32
     descriptor), embedding net on the matrix cores, contraction
     -> A [N, 4, E] (the reference's xyz_scater_a, dp_dp.py:344-415);
  2. fitting: DR = A^T A[:, :M2] -> fitting net -> Ei, and dEtot/dA;
  3. ``matpl_dp_fused_backward``: recomputed embedding forward + backward,
     dE/dRi -> force scatter and virial (dp_dp.py:417-469 semantics).
     With ``store=True`` the forward also keeps G(s) and dG/ds of every
     neighbor (``matpl_dp_fused_forward_store``) and the backward reads them
     (``matpl_dp_fused_backward_stored``) instead of running the net again.

Precision modes: "bf16" runs the embedding GEMMs on BF16 matrix cores with FP32
accumulation (and a fast exp2/rcp tanh); "fp32" runs them on FP32 matrix
cores with tanhf. Geometry, descriptors, contraction and scatter are FP32 in
both; the cutoff test, Etot and the virial are FP64.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib
from .calcops import _ptr, _stream
from .dp_fitting import FittingMFMA, segments, type_runs
from .dpnet import DPWeights

NVEC, NMAT, NTI = 7, 4, 16
TANH_K = 2.0 / np.log(2.0)      # 2 log2(e): BF16-mode pre-activation scale (see dp_fused.hip)
V_W1, V_B1, V_B2, V_B3, V_R2, V_R3, V_GPAD = range(NVEC)
M_W2F, M_W3F, M_W2B, M_W3B = range(NMAT)


def feat(r: int, h: int) -> int:
    """Feature held by accumulator register r in lane half h (32x32 MFMA layout)."""
    return 8 * (r // 4) + 4 * h + (r % 4)


def _emb_forward_np(layers, s: float) -> np.ndarray:
    """FP64 embedding net at one input (dp_embedding.py:73-94)."""
    x = np.array([s], np.float64)
    for W, b, r in layers:
        hdn = np.tanh(x @ W + (b if b is not None else 0.0))
        hs = hdn * r if r is not None else hdn
        if W.shape[1] == W.shape[0]:
            x = hs + x
        elif W.shape[1] == 2 * W.shape[0]:
            x = np.concatenate([x, x]) + hs
        else:
            x = hdn
    return x


@dataclass
class FusedWeights:
    vec: torch.Tensor      # [T*T, NVEC, 2, 16] float32
    mat: torch.Tensor      # [T*T, NMAT, 64, 16] bfloat16 or float32
    tinfo: torch.Tensor    # [T, NTI] float32
    E: int
    bf16: bool


def pack_fused(w: DPWeights, precision: str = "bf16", device="cuda") -> FusedWeights:
    cfg = w.cfg
    T = cfg.ntypes
    es = cfg.emb_sizes
    if len(es) != 3 or es[1] != es[0] or es[2] != es[0] or es[0] not in (16, 25, 32):
        raise ValueError("fused DP path needs embedding sizes [E, E, E] with E in {16, 25, 32}")
    E = es[0]
    bf16 = precision == "bf16"
    if precision not in ("bf16", "fp32"):
        raise ValueError("precision must be 'bf16' or 'fp32'")
    vec = np.zeros((T * T, NVEC, 2, 16), np.float32)
    mat = np.zeros((T * T, NMAT, 64, 16), np.float32)
    tinfo = np.zeros((T, NTI), np.float32)
    ri_pad = -w.davg / w.dstd                                         # Ri of an empty slot
    for ti in range(T):
        tinfo[ti, 0:4] = w.davg[ti]
        tinfo[ti, 4:8] = 1.0 / w.dstd[ti]
        tinfo[ti, 8:12] = w.dstd[ti]
        tinfo[ti, 12:16] = ri_pad[ti]
    for p in range(T * T):
        ti = p // T
        (W1, b1, _), (W2, b2, r2), (W3, b3, r3) = w.emb[p]
        gpad = _emb_forward_np(w.emb[p], float(ri_pad[ti, 0]))
        one = np.ones(E)
        vecs = {V_W1: W1[0], V_B1: b1 if b1 is not None else np.zeros(E), V_B2: b2 if b2 is not None else np.zeros(E),
                V_B3: b3 if b3 is not None else np.zeros(E), V_R2: r2 if r2 is not None else one,
                V_R3: r3 if r3 is not None else one, V_GPAD: gpad}
        for h in range(2):
            for r in range(16):
                f = feat(r, h)
                if f < E:
                    for k, v in vecs.items():
                        vec[p, k, h, r] = v[f]
        # the backward matrices carry the resnet scale of their layer's output feature f
        # (dh_in = dh_out + (W diag r) (dh_out (1 - t^2)))
        r2s, r3s = vecs[V_R2], vecs[V_R3]
        for lane in range(64):
            h, i = lane // 32, lane % 32
            for r in range(16):
                f = feat(r, h)
                if f < E and i < E:
                    mat[p, M_W2F, lane, r] = W2[f, i]
                    mat[p, M_W3F, lane, r] = W3[f, i]
                    mat[p, M_W2B, lane, r] = W2[i, f] * r2s[f]
                    mat[p, M_W3B, lane, r] = W3[i, f] * r3s[f]
    if bf16:
        # tanh(z) is evaluated as 1 - 2 / (2^y + 1) with y = TANH_K z: pre-scale the forward weights and
        # biases of every tanh layer (the backward matrices and the padding vector stay unscaled)
        for k in (V_W1, V_B1, V_B2, V_B3):
            vec[:, k] *= TANH_K
        mat[:, M_W2F] *= TANH_K
        mat[:, M_W3F] *= TANH_K
    mt = torch.as_tensor(mat, device=device)
    return FusedWeights(torch.as_tensor(vec, device=device), mt.to(torch.bfloat16) if bf16 else mt,
                        torch.as_tensor(tinfo, device=device), E, bf16)


@dataclass
class FusedTable:
    """A compressed embedding (matpl_amd.compress.CompressTable) in the fused kernels' layout."""
    tab: torch.Tensor      # [T*T, rows, 32, K] float32: pair, interval, feature (padded to 32), coefficient
    order: int
    rows: int
    n_fine: int
    x0: float
    dx: float


def table_value_np(ct, pair: int, s: float) -> np.ndarray:
    """G(s) of one pair from a CompressTable, in float64 (the lookup rule of CompressTable.lookup)."""
    f2, coef = ct.lookup(torch.tensor([s], dtype=torch.float64), pair)
    x = float(f2.reshape(-1)[0])
    c = coef[0].to(torch.float64).cpu().numpy()        # [E, K], highest power first
    v = c[:, 0].copy()
    for k in range(1, c.shape[1]):
        v = v * x + c[:, k]
    return v


def pack_table(ct, E: int, device) -> FusedTable:
    t = ct.table.detach().to(torch.float64).cpu().numpy()      # [P, rows, E, K]
    P, R, Et, K = t.shape
    if Et != E or E > 32:
        raise ValueError(f"compress table has width {Et}, the model {E}")
    if K != ct.order + 1:
        raise ValueError(f"compress table of order {ct.order} has {K} coefficients")
    pad = np.zeros((P, R, 32, K))
    pad[:, :, :E] = t
    return FusedTable(torch.as_tensor(pad, dtype=torch.float32, device=device), ct.order, R, ct.n_fine,
                      float(ct.x0), float(ct.dx))


class _Fitting:
    """Per-type fitting nets in FP32 torch with autograd for dE/dA."""

    def __init__(self, w: DPWeights, device, dtype=torch.float32):
        self.cfg = w.cfg
        self.nets = []
        for t in range(w.cfg.ntypes):
            self.nets.append([(torch.as_tensor(W, dtype=dtype, device=device),
                               torch.as_tensor(b, dtype=dtype, device=device) if b is not None else None,
                               torch.as_tensor(r, dtype=dtype, device=device) if r is not None else None)
                              for (W, b, r) in w.fit[t]])

    def __call__(self, A: torch.Tensor, type_map: torch.Tensor, type_ranges):
        """A [N, 4, E] -> (Ei [N], dEtot/dA [N, 4, E])."""
        M2 = self.cfg.M2
        A = A.detach().requires_grad_()
        Ei = torch.empty(A.shape[0], dtype=A.dtype, device=A.device)
        parts = []
        for t, a, b in segments(type_ranges):
            if b <= a:
                continue
            At = A[a:b]
            DR = (At.transpose(1, 2) @ At[:, :, :M2]).reshape(b - a, -1)
            x = DR
            layers = self.nets[t]
            for l, (W, bias, r) in enumerate(layers):
                z = x @ W + (bias if bias is not None else 0.0)
                if l == len(layers) - 1:
                    x = z
                    break
                hdn = torch.tanh(z)
                x = hdn if l == 0 else (hdn * r if r is not None else hdn) + x
            parts.append((a, b, x.reshape(-1)))
        Etot = sum(x.sum() for _, _, x in parts)
        (dA,) = torch.autograd.grad(Etot, A)
        for a, b, x in parts:
            Ei[a:b] = x.detach()
        return Ei, dA


def frame_types_of(type_map: torch.Tensor, atom_frame: Optional[torch.Tensor], nframes: int, ntypes: int):
    """[F, T] bool: the types that occur in each frame, i.e. the per-image atom_type list the
    reference's data loader hands DP.forward (dp_dp.py:150-165, 197)."""
    tm = type_map.reshape(-1).to(torch.int64)
    fr = torch.zeros_like(tm) if atom_frame is None else atom_frame.reshape(-1).to(torch.int64)
    occ = torch.zeros(nframes * ntypes, dtype=torch.bool, device=tm.device)
    occ[fr * ntypes + tm] = True
    return occ.view(nframes, ntypes)


def type_masks(type_present, nframes: int, ntypes: int) -> np.ndarray:
    """[F] uint32 bit masks from a [T] (every frame) or [F, T] presence table."""
    tp = np.asarray(type_present.cpu() if isinstance(type_present, torch.Tensor) else type_present).astype(bool)
    if tp.ndim == 1:
        tp = np.broadcast_to(tp, (nframes, tp.shape[0]))
    if tp.shape != (nframes, ntypes):
        raise ValueError(f"type_present must be [{ntypes}] or [{nframes}, {ntypes}], got {tp.shape}")
    masks = (tp.astype(np.uint64) << np.arange(ntypes, dtype=np.uint64)).sum(1).astype(np.uint32)
    if (masks == 0).any():
        raise ValueError("type_present: every frame needs at least one type in its list")
    return masks


class FusedDP:
    """DP E/F/V for one (or a batch of) frame(s) given a neighbor list.

    Frames are packed along the atom axis, each frame's atoms contiguous. The fitting runs over
    runs of atoms of one type (``type_ranges``: [(first, end)] per type, or [(type, first, end)]
    runs; default the runs of ``type_map``), so atoms grouped by type within each frame keep
    its tiles full.
    """

    def __init__(self, w: DPWeights, precision: str = "bf16", device="cuda", fitting: str = "mfma", compress=None,
                 store: bool = False):
        """fitting: "mfma" (hand-written matrix-core kernel) or "torch" (FP32 torch + autograd).
        compress: a matpl_amd.compress.CompressTable; the fused kernels then evaluate the embedding
        from the table (G and dG/ds by Horner) instead of the net, as DPTorch(compress=...) does.
        store: the forward keeps G(s) and dG/ds of every in-cutoff neighbor in a device buffer
        (fp16 pairs in BF16 mode, floats in FP32 mode) and the backward reads them instead of
        evaluating the embedding net a second time (net mode only)."""
        if store and compress is not None:
            raise ValueError("FusedDP: store mode applies to the embedding net, not to a compressed table")
        self.store = store
        self._gbuf = None
        self.w = w
        self.cfg = w.cfg
        self.device = torch.device(device)
        self.fw = pack_fused(w, precision, self.device)
        self.table = None
        if compress is not None:
            T = self.cfg.ntypes
            self.table = pack_table(compress, self.fw.E, self.device)
            # the empty-slot vector from the same table
            vec = self.fw.vec.cpu().numpy()
            ri_pad = -w.davg / w.dstd
            for p in range(T * T):
                g = table_value_np(compress, p, float(ri_pad[p // T, 0]))
                for h in range(2):
                    for r in range(16):
                        f = feat(r, h)
                        vec[p, V_GPAD, h, r] = g[f] if f < self.fw.E else 0.0
            self.fw.vec = torch.as_tensor(vec, device=self.device)
        self.fit = _Fitting(w, self.device)
        self.fitting = fitting
        self._mfma_fit = {}
        self.precision = precision
        self.events = None      # when a list: (start, end) HIP events around each backward launch
        self._masks = {}        # checked per-frame type masks on the device, by content and atoms

    def __call__(self, pos: torch.Tensor, lattice: torch.Tensor, type_map: torch.Tensor, nl_list: torch.Tensor,
                 nl_shift: torch.Tensor, nl_counts: torch.Tensor, mnl: int, atom_frame: Optional[torch.Tensor] = None,
                 frame_start: Optional[torch.Tensor] = None, nframes: int = 1, type_ranges=None,
                 rev: Optional[torch.Tensor] = None, type_present=None):
        """Returns (Ei [N], F [N,3] float32, V [nframes, 9] float64, status).

        With ``rev`` (neighbor.reverse_map) the neighbor forces are gathered
        without atomics; otherwise they are scattered with atomics.
        ``type_present`` ([T] or [nframes, T] bools, e.g. ``frame_types_of``) is each frame's
        atom_type list (dp_dp.py:150-165, 197): only its types' buckets take part and the
        descriptor is divided by m times their number (dp_dp.py:413). None: every model type
        (the LAMMPS pair style's use, where the list is the model's).
        """
        cfg = self.cfg
        N = pos.shape[0]
        T = cfg.ntypes
        dev = pos.device
        if type_ranges is None:
            type_ranges = type_runs(type_map.cpu().numpy())
        scale = 1.0 / (cfg.m * T)
        masks = None if type_present is None else self._type_masks(type_present, type_map, atom_frame, nframes)
        A = torch.empty((N, 4, self.fw.E), dtype=torch.float32, device=dev)
        status = torch.zeros(1, dtype=torch.int32, device=dev)
        head = (_ptr(pos), _ptr(lattice), _ptr(atom_frame), _ptr(frame_start), _ptr(masks), _ptr(nl_list),
                _ptr(nl_shift), _ptr(nl_counts), _ptr(type_map), _ptr(self.fw.vec))
        tail = (_ptr(self.fw.tinfo), N, T, mnl, cfg.m, self.fw.E, int(self.fw.bf16), float(cfg.rcut),
                float(cfg.rcut_smooth), scale)
        if self.table is None:
            args, sfx = head + (_ptr(self.fw.mat),) + tail, ""
        else:
            tb = self.table
            args, sfx = head + (_ptr(tb.tab), tb.order, tb.rows, tb.n_fine, tb.x0, tb.dx) + tail, "_compressed"
        gbuf = None
        if self.store:
            words = _lib.query("matpl_dp_fused_gstore_words", N, T, mnl, cfg.m, self.fw.E, int(self.fw.bf16))
            if words < 0:
                raise ValueError("FusedDP: no store layout for this embedding width")
            if self._gbuf is None or self._gbuf.numel() < words or self._gbuf.device != dev:
                self._gbuf = torch.empty(max(words, 4), dtype=torch.int32, device=dev)
            gbuf = self._gbuf
            _lib.call("matpl_dp_fused_forward_store", *args, _ptr(A), _ptr(gbuf), _ptr(status), _stream())
        else:
            _lib.call("matpl_dp_fused_forward" + sfx, *args, _ptr(A), _ptr(status), _stream())
        if self.fitting == "mfma":
            key = tuple(segments(type_ranges))
            if key not in self._mfma_fit:
                self._mfma_fit[key] = FittingMFMA(self.w, type_ranges, self.precision, dev)
            Ei, dA = self._mfma_fit[key](A)
        else:
            Ei, dA = self.fit(A, type_map, type_ranges)
        dA = dA.contiguous()
        V = torch.zeros((nframes, 9), dtype=torch.float64, device=dev)
        if rev is not None:
            F = torch.empty((N, 3), dtype=torch.float32, device=dev)
            fslot = torch.empty((N * T * mnl, 4), dtype=torch.float32, device=dev)
            fcenter = torch.empty((N, 3), dtype=torch.float32, device=dev)
        else:
            F = torch.zeros((N, 3), dtype=torch.float32, device=dev)
            fslot = fcenter = None
        if self.events is not None:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
        if gbuf is not None:
            _lib.call("matpl_dp_fused_backward_stored", *args, _ptr(dA), _ptr(gbuf), _ptr(rev), _ptr(fslot),
                      _ptr(fcenter), _ptr(F), _ptr(V), _ptr(status), _stream())
        else:
            _lib.call("matpl_dp_fused_backward" + sfx, *args, _ptr(dA), _ptr(rev), _ptr(fslot), _ptr(fcenter),
                      _ptr(F), _ptr(V), _ptr(status), _stream())
        if self.events is not None:
            e1.record()
            self.events.append((e0, e1))
        if rev is not None:
            _lib.call("matpl_dp_fused_gather", _ptr(nl_counts), _ptr(fslot), _ptr(fcenter), _ptr(F), T, mnl, N,
                      _stream())
        return Ei, F, V, status

    def _type_masks(self, type_present, type_map, atom_frame, nframes):
        """The per-frame type masks on the device. The first call with a given table and atom
        arrays checks that every type of a frame's atoms is in its list (the reference has no
        rule for atoms outside it); later calls reuse the checked tensor without a host sync."""
        masks = type_masks(type_present, nframes, self.cfg.ntypes)
        key = (masks.tobytes(), type_map.data_ptr(), None if atom_frame is None else atom_frame.data_ptr(),
               type_map.numel())
        t = self._masks.get(key)
        if t is None:
            occ = frame_types_of(type_map, atom_frame, nframes, self.cfg.ntypes).cpu().numpy()
            bits = (occ.astype(np.uint64) << np.arange(self.cfg.ntypes, dtype=np.uint64)).sum(1).astype(np.uint32)
            bad = np.nonzero(bits & ~masks)[0]
            if bad.size:
                raise ValueError(f"type_present: frame {int(bad[0])} has atoms of a type outside its list")
            t = torch.as_tensor(masks.view(np.int32), device=type_map.device)
            self._masks = {key: t}
        return t
