"""FP64 CPU oracle for the NEP path (TEST INFRASTRUCTURE).

The neighbor search is the C restatement in nep_oracle.c; the descriptor,
fitting net, energy and its derivatives are written here with torch on the
CPU in float64, gradients by autograd (a library routine) - the same split the
reference's CPU branch uses (src/model/nep_net.py:342-351, 536-603, 651-914).

  fk            nep_utilities.cuh:58-256 / nep_cpu.cpp:16-56 (fc, Chebyshev f_k)
  blm           nep_cpu.cpp:58-91 (24 real harmonics / r^l, as accumulate_s)
  q_from_s      nep_cpu.cpp:93-129 (C3B/C4B/C5B, nep_cpu.h:8-16)
  radial_feats  calculateNepFeat.cu:5-96 (loop stops at the first -1)
  angular_feats nep3_small_box.cuh:68-146
  nep_evaluate  NEP.forward + calculate_force_virial CPU branch (nep_net.py:293-648)

Pins (tests/test_oracle_nep.py): bit-exact neighbor lists and 1e-12 features
against oracle/_ref (the reference's own nep_cpu.cpp +
cpu_calculate_nepneighbor.cpp, compiled from source), Chebyshev closed form
T_k(x) = cos(k acos x), rotation invariance of the features, forces = -dE/dx
and virial = -dE/d(strain) by central differences.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import lib as _clib, _c, _f64p

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

# 3-body coefficients with the factor 2 of the m != 0 terms kept separate
# (q_nl = C3B[l0] s_l0^2 + 2 sum_{m>0} C3B[lm] s_lm^2), nep_cpu.h:8-13
C3B = np.array([
    0.238732414637843, 0.119366207318922, 0.119366207318922, 0.099471839432435, 0.596831036594608,
    0.596831036594608, 0.149207759148652, 0.149207759148652, 0.139260575205408, 0.104445431404056,
    0.104445431404056, 1.044454314040563, 1.044454314040563, 0.174075719006761, 0.174075719006761,
    0.011190581936149, 0.223811638722978, 0.223811638722978, 0.111905819361489, 0.111905819361489,
    1.566681471060845, 1.566681471060845, 0.195835183882606, 0.195835183882606])
C4B = np.array([-0.007499480826664, -0.134990654879954, 0.067495327439977, 0.404971964639861, -0.809943929279723])
C5B = np.array([0.026596810706114, 0.053193621412227, 0.026596810706114])
L_SLICES = [(0, 3), (3, 8), (8, 15), (15, 24)]


def _setup():
    L = _clib()
    if not hasattr(L, "_nep_ready"):
        L.oracle_nep_expand_box.argtypes = [_f64p, ctypes.c_double, _i64p, _f64p]
        L.oracle_nep_expand_box.restype = ctypes.c_double
        L.oracle_nep_neighbor.argtypes = [ctypes.c_int64, _i64p, _f64p, _f64p, _i64p, _f64p, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _i64p,
                                          _i64p, _i64p, _i64p, _f64p, _f64p]
        L.oracle_nep_neighbor.restype = ctypes.c_int
        L.oracle_nep_maxneigh.argtypes = [ctypes.c_int64, _i64p, _f64p, _f64p, _i64p, _f64p, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, _i64p, ctypes.c_int, _i64p, _i64p]
        L._nep_ready = True
    return L


def expand_box(lattice, cutoff):
    """(num_cell [3] int64, box [18], box_orig [9]) for one frame (lattice rows = vectors)."""
    lat = _c(np.asarray(lattice, np.float64).T.reshape(9), np.float64)
    nc = np.zeros(3, np.int64)
    box = np.zeros(18)
    _setup().oracle_nep_expand_box(lat, float(cutoff), nc, box)
    return nc, box, lat


def pack_frames(frames, cutoff_radial):
    """Packed NEP inputs for a list of Frames (positions Cartesian)."""
    nums = np.array([f.natoms for f in frames], np.int64)
    boxes, origs, ncs = [], [], []
    for f in frames:
        nc, box, lat = expand_box(f.lattice, cutoff_radial)
        boxes.append(box); origs.append(lat); ncs.append(nc)
    return dict(num_atoms=nums, box=np.concatenate(boxes), box_orig=np.concatenate(origs),
                num_cell=np.concatenate(ncs).astype(np.int64),
                position=np.concatenate([f.pos for f in frames]).astype(np.float64),
                atom_type_map=np.concatenate([f.types for f in frames]).astype(np.int64))


def neighbor(packed, rc_r, rc_a, maxr, maxa):
    L = _setup()
    n = int(packed["num_atoms"].sum())
    out = dict(NN_r=np.zeros(n, np.int64), NL_r=np.zeros((n, maxr), np.int64), NN_a=np.zeros(n, np.int64),
               NL_a=np.zeros((n, maxa), np.int64), Ri_r=np.zeros((n, maxr, 4)), Ri_a=np.zeros((n, maxa, 4)))
    err = L.oracle_nep_neighbor(len(packed["num_atoms"]), _c(packed["num_atoms"], np.int64),
                                _c(packed["box"], np.float64), _c(packed["box_orig"], np.float64),
                                _c(packed["num_cell"], np.int64), _c(packed["position"], np.float64), float(rc_r),
                                float(rc_a), maxr, maxa, n, out["NN_r"], out["NL_r"], out["NN_a"], out["NL_a"],
                                out["Ri_r"], out["Ri_a"])
    if err:
        raise OverflowError(f"max neighbors too small (atom {err - 1})")
    return out


def maxneigh(packed, rc_r, rc_a, ntypes, with_type):
    if with_type and np.asarray(packed["atom_type_map"]).max(initial=-1) >= ntypes:
        raise ValueError("maxneigh: atom type >= ntypes")
    L = _setup()
    n = int(packed["num_atoms"].sum())
    shape = (n, ntypes) if with_type else (n,)
    NN_r = np.zeros(shape, np.int64)
    NN_a = np.zeros(shape, np.int64)
    L.oracle_nep_maxneigh(len(packed["num_atoms"]), _c(packed["num_atoms"], np.int64), _c(packed["box"], np.float64),
                          _c(packed["box_orig"], np.float64), _c(packed["num_cell"], np.int64),
                          _c(packed["position"], np.float64), float(rc_r), float(rc_a), ntypes,
                          _c(packed["atom_type_map"], np.int64), int(with_type), NN_r, NN_a)
    return NN_r, NN_a


# ----------------------------------------------------------------- descriptor (torch FP64)
def fk(r: torch.Tensor, nb: int, rc: float) -> torch.Tensor:
    """f_k(r) = (T_k(x)+1)/2 fc(r), k < nb; 0 at and beyond rc."""
    inside = r < rc
    fc = torch.where(inside, 0.5 * torch.cos(math.pi * r / rc) + 0.5, torch.zeros_like(r))
    x = 2.0 * (r / rc - 1.0) ** 2 - 1.0
    T = [torch.ones_like(r), x]
    for k in range(2, nb):
        T.append(2.0 * x * T[-1] - T[-2])
    return torch.stack([(t + 1.0) * 0.5 * fc for t in T[:nb]], -1)


def blm(r: torch.Tensor, xyz: torch.Tensor) -> torch.Tensor:
    """The 24 real harmonic polynomials of (x,y,z)/r, in accumulate_s order."""
    u = xyz / r.unsqueeze(-1)
    x, y, z = u[..., 0], u[..., 1], u[..., 2]
    x2, y2, z2 = x * x, y * y, z * z
    xy2 = x2 - y2
    return torch.stack([
        z, x, y,
        3 * z2 - 1, x * z, y * z, xy2, 2 * x * y,
        (5 * z2 - 3) * z, (5 * z2 - 1) * x, (5 * z2 - 1) * y, xy2 * z, 2 * x * y * z, (x2 - 3 * y2) * x,
        (3 * x2 - y2) * y,
        (35 * z2 - 30) * z2 + 3, (7 * z2 - 3) * x * z, (7 * z2 - 3) * y * z, (7 * z2 - 1) * xy2,
        (7 * z2 - 1) * x * y * 2, (x2 - 3 * y2) * x * z, (3 * x2 - y2) * y * z, xy2 * xy2 - 4 * x2 * y2,
        4 * x * y * xy2], -1)


def q_from_s(s: torch.Tensor, l3: int, l4: int, l5: int) -> torch.Tensor:
    """s [N, na, 24] -> angular features [N, F3] in the [l][n] layout."""
    c = torch.as_tensor(C3B, dtype=s.dtype)
    w = torch.where(torch.as_tensor([k in (0, 3, 8, 15) for k in range(24)]), c, 2.0 * c)
    sq = w * s * s
    qs = [sq[..., a:b].sum(-1) for (a, b) in L_SLICES[:l3]]          # each [N, na]
    if l4 > 0:
        qs.append(C4B[0] * s[..., 3] ** 3 + C4B[1] * s[..., 3] * (s[..., 4] ** 2 + s[..., 5] ** 2) +
                  C4B[2] * s[..., 3] * (s[..., 6] ** 2 + s[..., 7] ** 2) +
                  C4B[3] * s[..., 6] * (s[..., 5] ** 2 - s[..., 4] ** 2) + C4B[4] * s[..., 4] * s[..., 5] * s[..., 7])
    if l5 > 0:
        s0 = s[..., 0] ** 2
        s12 = s[..., 1] ** 2 + s[..., 2] ** 2
        qs.append(C5B[0] * s0 * s0 + C5B[1] * s0 * s12 + C5B[2] * s12 * s12)
    return torch.cat(qs, -1)


def _valid(NL: torch.Tensor) -> torch.Tensor:
    """Slots before the first -1 (the reference loops stop there)."""
    return torch.cumprod((NL >= 0).to(torch.int64), -1).bool()


def radial_feats(c2, Ri, NL, atom_map, rc):
    c2 = torch.as_tensor(c2, dtype=torch.float64)
    NL = torch.as_tensor(NL)
    am = torch.as_tensor(atom_map)
    valid = _valid(NL)
    tj = am[NL.clamp(min=0)]
    ti = am.unsqueeze(1).expand_as(NL)
    f = fk(Ri[..., 0], c2.shape[-1], rc)                              # [N,S,nb]
    c = c2[ti, tj]                                                    # [N,S,nm,nb]
    return ((c * f.unsqueeze(-2)).sum(-1) * valid.unsqueeze(-1)).sum(1)


def angular_feats(c3, Ri, NL, atom_map, rc, l3, l4, l5):
    c3 = torch.as_tensor(c3, dtype=torch.float64)
    NL = torch.as_tensor(NL)
    am = torch.as_tensor(atom_map)
    valid = _valid(NL) & (Ri[..., 0] <= rc)
    tj = am[NL.clamp(min=0)]
    ti = am.unsqueeze(1).expand_as(NL)
    r = torch.where(valid, Ri[..., 0], torch.ones_like(Ri[..., 0]))
    f = fk(r, c3.shape[-1], rc)
    g = (c3[ti, tj] * f.unsqueeze(-2)).sum(-1) * valid.unsqueeze(-1)  # [N,S,na]
    b = blm(r, Ri[..., 1:4])                                          # [N,S,24]
    s = torch.einsum("isn,isl->inl", g, b)
    return q_from_s(s, l3, l4, l5)


def fit_forward(layers, x):
    for l, (W, b, r) in enumerate(layers):
        z = x @ torch.as_tensor(W) + (torch.as_tensor(b) if b is not None else 0.0)
        if l == len(layers) - 1:
            return z
        h = torch.tanh(z)
        x = h if l == 0 else (h * torch.as_tensor(r) if r is not None else h) + x
    return x


def nep_energy(w, Ri_r, NL_r, Ri_a, NL_a, atom_map):
    """Per-atom energies Ei [N] (torch, differentiable wrt Ri_r / Ri_a)."""
    cfg = w.cfg
    q = radial_feats(w.c2, Ri_r, NL_r, atom_map, cfg.rcut_radial)
    l3, l4, l5 = cfg.l_max
    if l3 > 0:
        q = torch.cat([q, angular_feats(w.c3, Ri_a, NL_a, atom_map, cfg.rcut_angular, l3, l4, l5)], -1)
    x = q * torch.as_tensor(w.q_scaler)
    am = torch.as_tensor(atom_map)
    Ei = torch.zeros(q.shape[0], dtype=torch.float64)
    for t in range(cfg.ntypes):
        m = am == t
        if m.any():
            Ei = Ei.index_put((m,), fit_forward(w.fit[t], x[m]).reshape(-1))
    return Ei, q


def nep_evaluate(frames, w, maxr=None, maxa=None):
    """E per frame, Ei, F [N,3], V per frame [F,9] (FP64), features, neighbor data."""
    cfg = w.cfg
    packed = pack_frames(frames, cfg.rcut_radial)
    NNr, NNa = maxneigh(packed, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes, False)
    maxr = maxr or int(NNr.max())
    maxa = maxa or int(NNa.max())
    nb = neighbor(packed, cfg.rcut_radial, cfg.rcut_angular, maxr, maxa)
    Ri_r = torch.as_tensor(nb["Ri_r"]).requires_grad_()
    Ri_a = torch.as_tensor(nb["Ri_a"]).requires_grad_()
    am = packed["atom_type_map"]
    Ei, q = nep_energy(w, Ri_r, nb["NL_r"], Ri_a, nb["NL_a"], am)
    gr, ga = torch.autograd.grad(Ei.sum(), [Ri_r, Ri_a], allow_unused=True)
    ga = torch.zeros_like(Ri_a) if ga is None else ga
    n = Ei.shape[0]
    F = np.zeros((n, 3))
    nf = len(frames)
    starts = np.concatenate([[0], np.cumsum(packed["num_atoms"])])
    V = np.zeros((nf, 9))
    frame_of = np.repeat(np.arange(nf), packed["num_atoms"])
    for Ri, g, NL in ((nb["Ri_r"], gr.numpy(), nb["NL_r"]), (nb["Ri_a"], ga.numpy(), nb["NL_a"])):
        valid = NL >= 0
        r = np.where(valid, Ri[..., 0], 1.0)
        # dE/d(dR) = dE/dr * dR/r + dE/d(xyz)   (the reference's Ri_d chain, nep_net.py:440-468)
        dEdR = g[..., 0:1] * Ri[..., 1:4] / r[..., None] + g[..., 1:4]
        dEdR = dEdR * valid[..., None]
        F += dEdR.sum(1)                                   # centre: -dE/dr_i = +sum dE/d(dR)
        ii, ss = np.nonzero(valid)
        np.add.at(F, NL[ii, ss], -dEdR[ii, ss])           # neighbor: -dE/dr_j
        for a in range(3):
            for c in range(3):
                np.add.at(V[:, 3 * a + c], frame_of[ii], -(Ri[ii, ss, 1 + a] * dEdR[ii, ss, c]))
    E = np.array([Ei[starts[k]:starts[k + 1]].sum().item() for k in range(nf)])
    return dict(E=E, Ei=Ei.detach().numpy(), F=F, V=V, q=q.detach().numpy(), packed=packed, nb=nb,
                dE_r=gr.numpy(), dE_a=ga.numpy())


def nep_efv_graph(frames, cfg, c2, c3, fit, q_scaler, maxr=None, maxa=None):
    """E [F], F [N,3], V [F,9] (FP64 torch on the CPU) as differentiable functions of the parameter
    leaves c2, c3 and fit[t] = [(W, b, r)], for force/virial-loss gradients by the double backward
    (the reference's training path, nep_net.py:503-648 with create_graph). Same conventions as
    nep_evaluate, whose numpy forces and virial the strain and position differences pin."""
    packed = pack_frames(frames, cfg.rcut_radial)
    NNr, NNa = maxneigh(packed, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes, False)
    nb = neighbor(packed, cfg.rcut_radial, cfg.rcut_angular, maxr or int(NNr.max()), maxa or int(NNa.max()))
    am = packed["atom_type_map"]
    Rr = torch.as_tensor(nb["Ri_r"]).clone().requires_grad_()
    Ra = torch.as_tensor(nb["Ri_a"]).clone().requires_grad_()
    l3, l4, l5 = cfg.l_max
    q = radial_feats(c2, Rr, nb["NL_r"], am, cfg.rcut_radial)
    if l3 > 0:
        q = torch.cat([q, angular_feats(c3, Ra, nb["NL_a"], am, cfg.rcut_angular, l3, l4, l5)], -1)
    x = q * torch.as_tensor(q_scaler)
    amt = torch.as_tensor(am)
    Ei = torch.zeros(q.shape[0], dtype=torch.float64)
    for t in range(cfg.ntypes):
        m = amt == t
        if m.any():
            Ei = Ei.index_put((m,), fit_forward(fit[t], x[m]).reshape(-1))
    nf = len(frames)
    frame = torch.repeat_interleave(torch.arange(nf), torch.as_tensor(packed["num_atoms"]))
    E = torch.zeros(nf, dtype=torch.float64).index_add(0, frame, Ei)
    gr, ga = torch.autograd.grad(E.sum(), [Rr, Ra], create_graph=True, allow_unused=True)
    F = torch.zeros((q.shape[0], 3), dtype=torch.float64)
    V = torch.zeros((nf, 9), dtype=torch.float64)
    for R, g, NL in ((Rr, gr, nb["NL_r"]), (Ra, ga, nb["NL_a"])):
        if g is None:
            continue
        NLt = torch.as_tensor(NL)
        valid = (NLt >= 0).unsqueeze(-1)
        Rd = R.detach()
        r = torch.where(valid[..., 0], Rd[..., 0], torch.ones_like(Rd[..., 0]))
        pair = (g[..., 0:1] * Rd[..., 1:4] / r.unsqueeze(-1) + g[..., 1:4]) * valid       # dE/d(dR)
        F = F + pair.sum(1)
        F = F.index_add(0, NLt.clamp(min=0).reshape(-1), -pair.reshape(-1, 3))
        outer = -(Rd[..., 1:4].unsqueeze(-1) * pair.unsqueeze(-2)).sum(1).reshape(-1, 9)   # per centre atom
        V = V.index_add(0, frame, outer)
    return E, F, V
