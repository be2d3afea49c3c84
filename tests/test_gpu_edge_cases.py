"""Edge cases of the hot path on the GPU against the FP64 oracle: an atom with no neighbor inside the
cutoff (only empty slots: the descriptor is the padding term alone and its force is zero), a frame of
two atoms, and the NEP fitting kernels with no atoms at all.

Tolerances as the main parity tests: fused DP FP32 within 1e-6 |E| and 1e-5 eV/A per force component;
NEP float64 within 1e-10 relative."""
import numpy as np
import pytest
import torch

import oracle
import oracle.nep as on
from matpl_amd import nep as mn
from matpl_amd import structures as st
from matpl_amd.dp_fused import FusedDP
from matpl_amd.dpnet import DPConfig, DPWeights
from matpl_amd.nepnet import NEPConfig, NEPWeights
from matpl_amd.neighbor import find_neighbore_gpu, reverse_map

pytestmark = pytest.mark.gpu


def _sparse_frame(n_cluster=12, box=24.0, seed=3):
    """A small cluster in a large cubic box plus one atom more than 8 A from every other atom and
    from their images."""
    rng = np.random.default_rng(seed)
    pts = [np.zeros(3)]
    while len(pts) < n_cluster:
        p = rng.uniform(0.0, 6.0, 3)
        if min(np.linalg.norm(p - q) for q in pts) > 2.0:
            pts.append(p)
    pts.append(np.array([box / 2 + 3.0, box / 2 + 3.0, box / 2 + 3.0]))
    pos = np.array(pts) + 1.0
    types = (np.arange(len(pts)) % 2).astype(np.int32)
    order = np.argsort(types, kind="stable")
    return st.Frame(np.eye(3) * box, pos[order], types[order], ["Hf", "O"])


def _fused_vs_oracle(f, dev):
    T = 2
    cfg = DPConfig(ntypes=T, m=60)
    w = DPWeights.random(cfg, seed=5, davg=np.tile([0.1, 0, 0, 0], (T, 1)), dstd=np.tile([0.2, 0.1, 0.1, 0.1], (T, 1)))
    ref = oracle.dp_evaluate(f, w)
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, T, cfg.m, cfg.rcut, device=dev, want_dR=False,
                            want_shift=True)
    L = torch.as_tensor(f.lattice, device=dev)
    pos = (nl.frac_wrapped @ L).contiguous()
    tm = torch.as_tensor(f.types, device=dev)
    rev, rst = reverse_map(nl, tm, cfg.m)
    assert int(rst.item()) == 0
    Ei, F, V, status = FusedDP(w, "fp32", dev)(pos, L.reshape(1, 9).contiguous(), tm, nl.list, nl.shift,
                                                nl.counts, cfg.m, rev=rev)
    torch.cuda.synchronize()
    assert int(status.item()) == 0
    Ei, F = Ei.double().cpu().numpy(), F.double().cpu().numpy()
    assert abs(Ei.sum() - ref["E"]) <= 1e-6 * abs(ref["E"])
    assert np.abs(F - ref["F"]).max() <= 1e-5
    return Ei, F, ref


def test_fused_isolated_atom(dev):
    f = _sparse_frame()
    print(f"case isolated atom: N={f.natoms}", flush=True)
    Ei, F, ref = _fused_vs_oracle(f, dev)
    lone = int(np.argmax(np.linalg.norm(f.pos - f.pos.mean(0), axis=1)))
    assert np.abs(F[lone]).max() == 0.0          # no neighbor within rcut: no force


def test_fused_two_atoms(dev):
    f = st.Frame(np.eye(3) * 15.0, np.array([[1.0, 1.0, 1.0], [3.1, 1.4, 0.8]]), np.array([0, 1], np.int32),
                 ["Hf", "O"])
    print("case two atoms", flush=True)
    _, F, _ = _fused_vs_oracle(f, dev)
    assert np.abs(F.sum(0)).max() <= 1e-6        # Newton's third law on the pair


def test_nep_isolated_atom(dev):
    f = _sparse_frame(seed=4)
    cfg = NEPConfig(ntypes=2, l_max=[4, 2, 1])
    w = NEPWeights.random(cfg, seed=13)
    print(f"case NEP isolated atom: N={f.natoms}", flush=True)
    want = on.nep_evaluate([f], w)
    gp = mn.pack_frames([f], cfg.rcut_radial, dev)
    nb = mn.neighbors(gp, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes)
    model = mn.NEPTorch(w, torch.float64, dev)
    Etot, Ei, F, V = model(nb["NN_r"], nb["NL_r"], nb["Ri_r"], nb["NN_a"], nb["NL_a"], nb["Ri_a"], gp["num_atoms"],
                           gp["atom_type_map"], create_graph=False)
    rel = lambda a, b: float(np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-30))
    assert rel(Ei.detach().cpu().numpy(), want["Ei"]) < 1e-10
    assert rel(F.detach().cpu().numpy(), want["F"]) < 1e-10


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_nep_fit_kernels_no_atoms(dev, dtype):
    D, H, T = 30, 30, 2
    q = torch.zeros((0, D), dtype=dtype, device=dev, requires_grad=True)
    qs = torch.ones(D, dtype=dtype, device=dev)
    W0 = torch.randn((T, D, H), dtype=dtype, device=dev, requires_grad=True)
    b0 = torch.randn((T, H), dtype=dtype, device=dev, requires_grad=True)
    w1 = torch.randn((T, H), dtype=dtype, device=dev, requires_grad=True)
    b1 = torch.randn(T, dtype=dtype, device=dev, requires_grad=True)
    empty = torch.zeros(0, dtype=torch.int64, device=dev)
    plan = mn.FitPlan([empty, empty], dtype == torch.float64, D, H, dev)
    Ei, g = mn._NepFitFn.apply(q, qs, W0, b0, w1, b1, plan)
    assert Ei.shape == (0,) and g.shape == (0, D)
    grads = torch.autograd.grad(Ei.sum() + g.sum(), [q, W0, b0, w1, b1])
    assert grads[0].shape == (0, D)
    assert all(float(x.abs().max()) == 0.0 for x in grads[1:])   # no atoms: zero parameter gradients


def test_nep_list_widths_reused_and_refreshed(dev):
    """mn.neighbors keeps the list widths of a batch: a second call builds the same lists without
    the counting pass, and after the atoms move closer (more neighbors than the kept widths) the
    overflow makes it count again and widen the lists."""
    from matpl_amd import structures as st2
    frames = st2.agausic_frames(nframes=2, nmin=40, nmax=60, seed=5)
    cfg = NEPConfig(ntypes=4, l_max=[4, 2, 0])
    gp = mn.pack_frames(frames, cfg.rcut_radial, dev)
    nb1 = mn.neighbors(gp, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes)
    nb2 = mn.neighbors(gp, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes)
    for k in ("NN_r", "NN_a", "NL_r", "NL_a", "Ri_r", "Ri_a"):
        assert torch.equal(nb1[k], nb2[k]), k
    w = nb1["NL_r"].shape[1], nb1["NL_a"].shape[1]
    print(f"case list widths {w}", flush=True)
    # compress every frame about its first atom: more neighbors within the cutoffs
    pos = gp["position"]
    start = 0
    for f in frames:
        c = pos[start:start + f.natoms].mean(0)
        pos[start:start + f.natoms] = c + 0.7 * (pos[start:start + f.natoms] - c)
        start += f.natoms
    nb3 = mn.neighbors(gp, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes)
    fresh = dict(gp)
    fresh.pop(("_nep_list_widths", float(cfg.rcut_radial), float(cfg.rcut_angular)))
    nb4 = mn.neighbors(fresh, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes)
    assert int(nb4["NN_r"].max()) > w[0] or int(nb4["NN_a"].max()) > w[1]   # the kept widths overflowed
    for k in ("NN_r", "NN_a", "NL_r", "NL_a", "Ri_r", "Ri_a"):
        assert torch.equal(nb3[k], nb4[k]), k
