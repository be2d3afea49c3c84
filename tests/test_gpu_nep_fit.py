"""NEP fitting kernels (matpl_nep_fit_forward / _backward) against torch autograd of the same nets.

The reference's fitting net (src/model/nep_fitting.py:40-163, network_size [H, 1]) is plain torch,
so a float64 torch restatement of it is the parity reference here: Ei = tanh(x W0 + b0) w1 + b1 with
x = q_scaler * q, per atom type. Tolerances: float64 1e-12 relative (max over entries), float32 2e-5.
The backward is checked for L = sum gE Ei + sum gg . dEi/dq, i.e. through the first derivative of
the net (torch double backward), which is what a force loss sends into it.
"""
import numpy as np
import pytest
import torch

from matpl_amd import nep as mn

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).abs().max() / max(float(b.abs().max()), 1e-30))


def _torch_ref(q, qs, W0, b0, w1, b1, types):
    """(Ei, g) from float64 torch with graph, per type."""
    x = q * qs
    Ei = torch.zeros(q.shape[0], dtype=q.dtype, device=q.device)
    for t in range(W0.shape[0]):
        m = types == t
        if m.any():
            h = torch.tanh(x[m] @ W0[t] + b0[t])
            Ei = Ei.index_put((m,), h @ w1[t] + b1[t])
    (g,) = torch.autograd.grad(Ei.sum(), q, create_graph=True)
    return Ei, g


def _case(dev, dtype, N, T, D, H, seed, empty_type=None):
    rng = np.random.default_rng(seed)
    types = rng.integers(0, T, size=N)
    if empty_type is not None:
        types[types == empty_type] = (empty_type + 1) % T
    mk = lambda *s, sc=1.0: torch.tensor(rng.normal(size=s) * sc, dtype=torch.float64, device=dev)
    q = mk(N, D, sc=0.7).requires_grad_()
    qs = torch.tensor(rng.uniform(0.5, 2.0, size=D), dtype=torch.float64, device=dev)
    W0 = mk(T, D, H, sc=1.0 / np.sqrt(D)).requires_grad_()
    b0 = mk(T, H).requires_grad_()
    w1 = mk(T, H, sc=1.0 / np.sqrt(H)).requires_grad_()
    b1 = mk(T).requires_grad_()
    tm = torch.as_tensor(types, device=dev)
    return q, qs, W0, b0, w1, b1, tm


CASES = [(1000, 3, 30, 30, 1, None), (257, 4, 35, 100, 2, 2), (65, 1, 25, 40, 3, None), (33, 2, 30, 30, 4, 0)]


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-5)])
@pytest.mark.parametrize("N,T,D,H,seed,empty", CASES)
def test_fit_kernels_vs_torch_double_backward(dev, dtype, tol, N, T, D, H, seed, empty):
    print(f"case nep fit N={N} T={T} D={D} H={H} {dtype} empty type {empty}", flush=True)
    q, qs, W0, b0, w1, b1, tm = _case(dev, dtype, N, T, D, H, seed, empty)
    Ei0, g0 = _torch_ref(q, qs, W0, b0, w1, b1, tm)
    rng = np.random.default_rng(seed + 100)
    gE = torch.tensor(rng.normal(size=N), dtype=torch.float64, device=dev)
    gg = torch.tensor(rng.normal(size=(N, D)), dtype=torch.float64, device=dev)
    want = torch.autograd.grad((Ei0 * gE).sum() + (g0 * gg).sum(), [q, W0, b0, w1, b1])

    leaves = [x.detach().to(dtype).requires_grad_() for x in (q, W0, b0, w1, b1)]
    qd, W0d, b0d, w1d, b1d = leaves
    tidx = [torch.nonzero(tm == t).reshape(-1) for t in range(T)]
    plan = mn.FitPlan(tidx, dtype == torch.float64, D, H, dev)
    Ei, g = mn._NepFitFn.apply(qd, qs.to(dtype), W0d, b0d, w1d, b1d, plan)
    torch.cuda.synchronize()
    assert _rel(Ei, Ei0) < tol
    assert _rel(g, g0) < tol
    got = torch.autograd.grad((Ei * gE.to(dtype)).sum() + (g * gg.to(dtype)).sum(), leaves)
    for name, a, b in zip(("dq", "dW0", "db0", "dw1", "db1"), got, want):
        if empty is not None and name != "dq":
            assert float(a[empty].abs().max()) == 0.0, name      # the empty type's rows stay zero
        assert _rel(a, b) < 10 * tol, (name, _rel(a, b))


def test_fit_kernels_are_deterministic(dev):
    """Two backward calls give bit-identical parameter gradients (per-tile rows added in tile order)."""
    q, qs, W0, b0, w1, b1, tm = _case(dev, torch.float64, 3000, 2, 30, 30, 9)
    tidx = [torch.nonzero(tm == t).reshape(-1) for t in range(2)]
    plan = mn.FitPlan(tidx, True, 30, 30, dev)
    outs = []
    for _ in range(2):
        Ei, g = mn._NepFitFn.apply(q, qs, W0, b0, w1, b1, plan)
        outs.append(torch.autograd.grad(Ei.sum() + (g ** 2).sum(), [W0, b0, w1, b1]))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 1e-4)])
def test_model_kernel_path_equals_torch_layers(dev, dtype, tol):
    """NEPTorch with the fitting kernels vs the same model on per-type torch layers: energies, forces,
    virials and the E+F+V-loss gradient of every parameter (c2, c3, nets), 4-type Ag-Au-Si-C frames."""
    from matpl_amd import structures as st
    from matpl_amd.nepnet import NEPConfig, NEPWeights
    frames = st.agausic_frames(nframes=3, nmin=40, nmax=90, seed=8)
    cfg = NEPConfig(ntypes=4, l_max=[4, 2, 1])
    w = NEPWeights.random(cfg, seed=21)
    gp = mn.pack_frames(frames, cfg.rcut_radial, dev)
    nb = mn.neighbors(gp, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes)
    res = []
    for kernel in (True, False):
        model = mn.NEPTorch(w, dtype, dev)
        assert model.fit_kernel
        model.fit_kernel = kernel
        E, Ei, F, V = model(nb["NN_r"], nb["NL_r"], nb["Ri_r"], nb["NN_a"], nb["NL_a"], nb["Ri_a"], gp["num_atoms"],
                            gp["atom_type_map"], create_graph=True)
        loss = (E ** 2).sum() + (F ** 2).sum() + (V ** 2).sum()
        params = [p for p in model.parameters() if p.requires_grad]
        res.append((E, F, V, torch.autograd.grad(loss, params)))
    (E1, F1, V1, g1), (E2, F2, V2, g2) = res
    assert _rel(E1, E2) < tol and _rel(F1, F2) < tol and _rel(V1, V2) < tol
    for a, b in zip(g1, g2):
        assert _rel(a, b) < 10 * tol
