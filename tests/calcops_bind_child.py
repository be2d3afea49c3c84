"""Child process of test_gpu_calcops_bind.py: loads libCalcOps_bind.so the way the reference's
models load their operator library and runs all ten CalcOps_cuda operators against the FP64
oracle, without the Python registration (matpl_amd.calcops is never imported).

The library is reached through a reference-shaped tree, <tmp>/src/op/build/lib/libCalcOps_bind.so,
from a module at <tmp>/src/model/ whose load line is the reference's (src/model/dp_dp.py:13-20,
src/model/nep_net.py:17-24). Prints one JSON line of errors."""
import importlib.util
import json
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import oracle                                                            # noqa: E402
import oracle.nep as on                                                  # noqa: E402
from oracle.compress import horner                                       # noqa: E402
from matpl_amd import structures as st                                   # noqa: E402
from matpl_amd.nepnet import NEPConfig, NEPWeights                       # noqa: E402

LIB = os.path.join(ROOT, "matpl_amd", "libCalcOps_bind.so")
LOADER = '''import os
import torch
lib_path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "op/build/lib/libCalcOps_bind.so")
torch.ops.load_library(lib_path)
CalcOps = torch.ops.CalcOps_cuda
'''


def load(tmp):
    libdir = os.path.join(tmp, "src", "op", "build", "lib")
    os.makedirs(libdir, exist_ok=True)
    os.makedirs(os.path.join(tmp, "src", "model"), exist_ok=True)
    os.symlink(LIB, os.path.join(libdir, "libCalcOps_bind.so"))
    path = os.path.join(tmp, "src", "model", "ops_loader.py")
    with open(path, "w") as f:
        f.write(LOADER)
    spec = importlib.util.spec_from_file_location("ops_loader", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.CalcOps


def rel(a, b):
    a, b = (x.detach().cpu() if isinstance(x, torch.Tensor) else x for x in (a, b))
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max(initial=0.0) / max(np.abs(b).max(initial=0.0), 1e-30))


def ri_d(Ri):
    """d[r, x, y, z]/d(dR) (nep_net.py:440-468): row 0 = dR / r, rows 1-3 the identity; 0 for empty slots."""
    r = Ri[..., 0:1]
    ok = (r.abs() > 1e-5).to(Ri.dtype)
    row0 = Ri[..., 1:4] / torch.where(r.abs() > 1e-5, r, torch.ones_like(r)) * ok
    eye = torch.eye(3, dtype=Ri.dtype, device=Ri.device).expand(Ri.shape[0], Ri.shape[1], 3, 3) * ok.unsqueeze(-1)
    return torch.cat([row0.unsqueeze(-2), eye], -2).contiguous()


def fit_forward(layers, x):
    for l, (W, b, r) in enumerate(layers):
        z = x @ W + b
        if l == len(layers) - 1:
            return z
        h = torch.tanh(z)
        x = h if l == 0 else (h * r if r is not None else h) + x
    return x


def nep_step(ops, w, dev, dtype, frames, create_graph):
    """NEP E/F/V through the ten-op library only, as the reference's GPU branch chains them."""
    cfg = w.cfg
    pk = on.pack_frames(frames, cfg.rcut_radial)
    T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    num, box, bo, nc = T(pk["num_atoms"], torch.int64), T(pk["box"], torch.float64), \
        T(pk["box_orig"], torch.float64), T(pk["num_cell"], torch.int64)
    pos, am = T(pk["position"], torch.float64), T(pk["atom_type_map"], torch.int64)
    NNr, NNa = ops.calculate_maxneigh(num, box, bo, nc, pos, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes, am, False)
    maxr, maxa = int(NNr.max().item()), int(NNa.max().item())
    NN_r, NN_a, NL_r, NL_a, Ri_r, Ri_a = ops.calculate_neighbor(num, am, am, box, bo, nc, pos, cfg.rcut_radial,
                                                               cfg.rcut_angular, maxr, maxa, True)
    params = dict(c2=T(w.c2, dtype).requires_grad_(), c3=T(w.c3, dtype).requires_grad_())
    fit = [[(T(W, dtype).requires_grad_(), T(b, dtype).requires_grad_(),
             None if r is None else T(r, dtype).requires_grad_()) for (W, b, r) in layers] for layers in w.fit]
    Rr = Ri_r.to(dtype).requires_grad_()
    Ra = Ri_a.to(dtype).requires_grad_()
    l3, l4, l5 = cfg.l_max
    e = torch.empty(0, dtype=dtype, device=dev)
    q = torch.cat([ops.calculateNepFeat(params["c2"], Rr, NL_r, am, e, cfg.rcut_radial, 0)[0],
                   ops.calculateNepMbFeat(params["c3"], Ra, NL_a, am, e, cfg.nr, l3, l4, l5, cfg.rcut_angular)[0]], -1)
    x = q * T(w.q_scaler, dtype)
    Ei = torch.zeros(q.shape[0], dtype=dtype, device=dev)
    for t in range(cfg.ntypes):
        m = torch.nonzero(am == t).reshape(-1)
        Ei = Ei.index_put((m,), fit_forward(fit[t], x[m]).reshape(-1))
    frame = torch.repeat_interleave(torch.arange(len(frames), device=dev), num)
    Etot = torch.zeros(len(frames), dtype=dtype, device=dev).index_add(0, frame, Ei)
    gr, ga = torch.autograd.grad(Etot.sum(), [Rr, Ra], create_graph=create_graph)
    F = V = None
    for R, g, NL in ((Rr, gr, NL_r), (Ra, ga, NL_a)):
        d = ri_d(R.detach())
        centre = -(g.reshape(g.shape[0], 1, -1) @ d.reshape(d.shape[0], -1, 3)).squeeze(1)   # nep_net.py:608
        f = ops.calculateNepForce(NL, g, d, centre)[0]
        v = ops.calculateNepVirial(NL, g, R.detach()[..., 1:].contiguous(), d, num)[0]
        F = f if F is None else F + f
        V = v if V is None else V + v
    plist = [params["c2"], params["c3"]] + [p for layers in fit for (W, b, r) in layers for p in (W, b, r)
                                             if p is not None]
    return dict(E=Etot, F=-F, V=-V, q=q, NL_r=NL_r, NL_a=NL_a, Ri_r=Ri_r, Ri_a=Ri_a, NN_r=NN_r, NN_a=NN_a,
                params=plist, pk=pk, maxr=maxr, maxa=maxa)


def oracle_force_loss_grads(w, frames, maxr, maxa):
    """d(sum F^2 + sum E^2)/d(c2, c3, fitting parameters) by the oracle's FP64 torch double backward."""
    cfg = w.cfg
    pk = on.pack_frames(frames, cfg.rcut_radial)
    nb = on.neighbor(pk, cfg.rcut_radial, cfg.rcut_angular, maxr, maxa)
    am = pk["atom_type_map"]
    c2 = torch.tensor(w.c2, requires_grad=True)
    c3 = torch.tensor(w.c3, requires_grad=True)
    Rr = torch.as_tensor(nb["Ri_r"]).clone().requires_grad_()
    Ra = torch.as_tensor(nb["Ri_a"]).clone().requires_grad_()
    q = torch.cat([on.radial_feats(c2, Rr, nb["NL_r"], am, cfg.rcut_radial),
                   on.angular_feats(c3, Ra, nb["NL_a"], am, cfg.rcut_angular, *cfg.l_max)], -1)
    fit = [[(torch.tensor(W, requires_grad=True), torch.tensor(b, requires_grad=True),
             None if r is None else torch.tensor(r, requires_grad=True)) for (W, b, r) in layers] for layers in w.fit]
    x = q * torch.as_tensor(w.q_scaler)
    amt = torch.as_tensor(am)
    Ei = torch.zeros(q.shape[0], dtype=torch.float64)
    for t in range(cfg.ntypes):
        m = amt == t
        Ei = Ei.index_put((m,), on.fit_forward(fit[t], x[m]).reshape(-1))
    frame = torch.repeat_interleave(torch.arange(len(frames)), torch.as_tensor(pk["num_atoms"]))
    Etot = torch.zeros(len(frames), dtype=torch.float64).index_add(0, frame, Ei)
    gr, ga = torch.autograd.grad(Etot.sum(), [Rr, Ra], create_graph=True)
    F = 0.0
    for R, g, NL in ((Rr, gr, nb["NL_r"]), (Ra, ga, nb["NL_a"])):
        NLt = torch.as_tensor(NL)
        valid = (NLt >= 0).unsqueeze(-1)
        r = torch.where(valid[..., 0], R[..., 0].detach(), torch.ones_like(R[..., 0]))
        pair = (g[..., 0:1] * R[..., 1:4].detach() / r.unsqueeze(-1) + g[..., 1:4]) * valid
        F = F + pair.sum(1).index_add(0, NLt.clamp(min=0).reshape(-1), -pair.reshape(-1, 3))
    loss = (F ** 2).sum() + (Etot ** 2).sum()
    params = [c2, c3] + [p for layers in fit for (W, b, r) in layers for p in (W, b, r) if p is not None]
    return loss.item(), torch.autograd.grad(loss, params)


def main(tmp):
    dev = torch.device("cuda:0")
    ops = load(tmp)
    out = {}
    # ---------------------------------------------------------------- NEP
    frames = [st.zincblende_sic((2, 2, 2), seed=4), st.agausic_frames(nframes=2, nmin=30, nmax=50, seed=5)[0]]
    cfg = NEPConfig(ntypes=4, l_max=[4, 2, 1])
    w = NEPWeights.random(cfg, seed=19)
    r = nep_step(ops, w, dev, torch.float64, frames, create_graph=True)
    want = on.nep_evaluate(frames, w, r["maxr"], r["maxa"])
    nb = want["nb"]
    out["nl_exact"] = bool(all(np.array_equal(r[k].detach().cpu().numpy(), nb[k])
                               for k in ("NN_r", "NN_a", "NL_r", "NL_a", "Ri_r", "Ri_a")))
    mr, ma = on.maxneigh(r["pk"], cfg.rcut_radial, cfg.rcut_angular, 4, True)
    pk = r["pk"]
    T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    g = ops.calculate_maxneigh(T(pk["num_atoms"], torch.int64), T(pk["box"], torch.float64),
                               T(pk["box_orig"], torch.float64), T(pk["num_cell"], torch.int64),
                               T(pk["position"], torch.float64), cfg.rcut_radial, cfg.rcut_angular, 4,
                               T(pk["atom_type_map"], torch.int64), True)
    out["maxneigh_exact"] = bool(np.array_equal(g[0].cpu().numpy(), mr) and np.array_equal(g[1].cpu().numpy(), ma))
    out["dq"] = rel(r["q"].detach().cpu(), want["q"])
    out["dE"] = rel(r["E"].detach().cpu(), want["E"])
    out["dF"] = rel(r["F"].detach().cpu(), want["F"])
    out["dV"] = rel(r["V"].detach().cpu(), want["V"])
    q = ops.calculate_descriptor(T(w.c2, torch.float64), T(w.c3, torch.float64), r["Ri_r"], r["NL_r"],
                                 T(pk["atom_type_map"], torch.int64), cfg.rcut_radial, cfg.rcut_angular, r["maxr"],
                                 *cfg.l_max)[0]
    out["d_descriptor"] = rel(q.cpu(), want["q"])          # the angular part over the radial list
    loss = (r["F"] ** 2).sum() + (r["E"] ** 2).sum()
    got = torch.autograd.grad(loss, r["params"])
    l_or, want_g = oracle_force_loss_grads(w, frames, r["maxr"], r["maxa"])
    out["d_loss"] = abs(loss.item() - l_or) / abs(l_or)
    out["d_grad"] = max(rel(a.detach().cpu(), b) for a, b in zip(got, want_g))
    # float32 through the same library
    r32 = nep_step(ops, w, dev, torch.float32, frames, create_graph=False)
    out["dF_fp32_rmse"] = float(np.sqrt(np.mean((r32["F"].detach().cpu().double().numpy() - want["F"]) ** 2)))
    # ---------------------------------------------------------------- DP
    rng = np.random.default_rng(3)
    B, N, S, ng = 1, 40, 24, 5
    nl = rng.integers(-1, N + ng, size=(B, N, S)).astype(np.int32)
    dE = rng.normal(size=(B, N, S * 4))
    Rid = rng.normal(size=(B, N, S * 4, 3))
    Rij = rng.normal(size=(B, N, S, 3))
    Fc = rng.normal(size=(B, N, 3))
    Tt = lambda a: torch.as_tensor(a, device=dev)
    f = ops.calculateForce(Tt(nl).unsqueeze(2), Tt(dE).unsqueeze(2), Tt(Rid), Tt(Fc), torch.tensor(ng))[0]
    out["dp_force"] = rel(f.cpu(), oracle.dp_force_op(nl, dE, Rid, Fc, ng))
    v, av = ops.calculateVirial(Tt(nl).unsqueeze(2), Tt(dE).unsqueeze(2), Tt(Rij), Tt(Rid), torch.tensor(ng))
    V0, AV0 = oracle.dp_virial_op(nl, dE, Rij, Rid, ng)
    out["dp_virial"] = max(rel(v.cpu(), V0), rel(av.cpu(), AV0))
    x = rng.uniform(0, 0.01, size=(30, 1))
    coef = rng.normal(size=(30, 25, 4))
    G = ops.calculateCompress(Tt(x), Tt(coef))[0].cpu().numpy()
    out["dp_compress"] = rel(G, horner(x, coef))
    out["python_ops_imported"] = "matpl_amd.calcops" in sys.modules or "matpl_amd.nep_ops" in sys.modules
    print(json.dumps(out))


if __name__ == "__main__":
    main(sys.argv[1])
