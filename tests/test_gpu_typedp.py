"""Type-embedding DP on the GPU: the reference's form (shared embedding net on [s, v(t_j)], TypeDPTorch)
equals the folded per-pair DP path and the FP64 oracle on the folded weights; the fused BF16 kernels
run the folded model within the BF16 force bar; a force-loss gradient with respect to a
type-embedding weight matches a central difference of the oracle's loss."""
import copy

import numpy as np
import pytest
import torch

import oracle
from matpl_amd import structures as st
from matpl_amd.dp import DPTorch
from matpl_amd.dp_fused import FusedDP
from matpl_amd.dpnet import DPConfig
from matpl_amd.neighbor import find_neighbore_gpu, image_dr, reverse_map
from matpl_amd.typedp import TypeDPTorch, TypeDPWeights, fold, type_vectors

pytestmark = pytest.mark.gpu


def _setup():
    f = st.fluorite_hfo2((2, 2, 2), seed=9)
    cfg = DPConfig(ntypes=2, m=100)
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 2, cfg.m, cfg.rcut)
    davg, dstd = oracle.dp_calc_stat(lst, dR, f.types, 2, cfg.rcut_smooth, cfg.rcut)
    tv = type_vectors([72, 8], ["atomic_number", "q"], values={"q": {72: 1.3, 8: 3.44}}, ranges={"q": (0.7, 4.0)})
    return f, cfg, TypeDPWeights.random(cfg, tv, seed=4, davg=davg, dstd=dstd)


def _eval(model, f, dev):
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, model.cfg.m, model.cfg.rcut, device=dev)
    return model(nl.list.unsqueeze(0), torch.as_tensor(f.types, device=dev), image_dr(nl.list, nl.dR).unsqueeze(0))


def test_reference_form_equals_folded_and_oracle(dev):
    f, cfg, tw = _setup()
    print(f"case typedp N={f.natoms}", flush=True)
    E1, Ei1, F1, V1 = _eval(TypeDPTorch(tw, torch.float64, dev), f, dev)
    E2, Ei2, F2, V2 = _eval(DPTorch(fold(tw), torch.float64, dev), f, dev)
    ref = oracle.dp_evaluate(f, fold(tw))
    assert abs(E1.item() - E2.item()) < 1e-12 * abs(E2.item())
    assert torch.allclose(F1, F2, rtol=1e-12, atol=1e-12) and torch.allclose(V1, V2, rtol=1e-12, atol=1e-10)
    assert abs(E1.item() - ref["E"]) < 1e-10 * abs(ref["E"])
    assert np.abs(F1[0].detach().cpu().numpy() - ref["F"]).max() < 1e-9


def test_fused_bf16_runs_folded_model(dev):
    f, cfg, tw = _setup()
    w = fold(tw)
    ref = oracle.dp_evaluate(f, w)
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, cfg.m, cfg.rcut, device=dev, want_dR=False,
                            want_shift=True)
    L = torch.as_tensor(f.lattice, device=dev)
    tm = torch.as_tensor(f.types, device=dev)
    rev, rst = reverse_map(nl, tm, cfg.m)
    Ei, F, V, status = FusedDP(w, "bf16", dev)((nl.frac_wrapped @ L).contiguous(), L.reshape(1, 9).contiguous(), tm,
                                               nl.list, nl.shift, nl.counts, cfg.m, rev=rev)
    torch.cuda.synchronize()
    assert int(status.item()) == 0 and int(rst.item()) == 0
    rmse = float(np.sqrt(np.mean((F.double().cpu().numpy() - ref["F"]) ** 2)))
    assert rmse < 1e-3, rmse


def test_type_weight_gradient_vs_finite_difference(dev):
    f, cfg, tw = _setup()
    model = TypeDPTorch(tw, torch.float64, dev)
    E, Ei, F, V = _eval(model, f, dev)
    loss = (F ** 2).sum() + (E ** 2).sum()
    W1 = model.shared.W[0]
    (g,) = torch.autograd.grad(loss, [W1])

    def oracle_loss(tw2):
        r = oracle.dp_evaluate(f, fold(tw2))
        return float((r["F"] ** 2).sum() + r["E"] ** 2)

    h = 1e-5
    for idx in ((1, 3), (2, 17), (0, 5)):       # rows 1, 2: type-vector weights; row 0: the s weight
        vals = []
        for sgn in (1, -1):
            tw2 = copy.deepcopy(tw)
            W, b, r = tw2.emb[0]
            W = W.copy()
            W[idx] += sgn * h
            tw2.emb[0] = (W, b, r)
            vals.append(oracle_loss(tw2))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(g[idx].item() - fd) <= 1e-6 * max(1.0, abs(fd)), (idx, g[idx].item(), fd)


def test_default_typedp_checkpoint_vs_oracle(dev, tmp_path):
    """A default-config TypeDP model (five physical properties from the element table,
    nn_feature_type.py:35) saved in MatPL's checkpoint layout and read back: the reference form on the
    HIP ops (TypeDPTorch) and the fused BF16 kernels on the folded model against the FP64 oracle."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_dp_ckpt import _typedp_checkpoint
    from matpl_amd.dp_ckpt import read_dp_checkpoint
    path = str(tmp_path / "typedp.ckpt")
    _typedp_checkpoint(path)
    w, order, info = read_dp_checkpoint(path)
    f = st.fluorite_hfo2((2, 2, 2), seed=9)
    print(f"case default TypeDP checkpoint N={f.natoms}", flush=True)
    ref = oracle.dp_evaluate(f, w)
    E1, Ei1, F1, V1 = _eval(TypeDPTorch(info["typedp"], torch.float64, dev), f, dev)
    assert abs(E1.item() - ref["E"]) < 1e-10 * abs(ref["E"])
    assert np.abs(F1[0].detach().cpu().numpy() - ref["F"]).max() < 1e-9
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, w.cfg.m, w.cfg.rcut, device=dev, want_dR=False,
                            want_shift=True)
    L = torch.as_tensor(f.lattice, device=dev)
    tm = torch.as_tensor(f.types, device=dev)
    rev, rst = reverse_map(nl, tm, w.cfg.m)
    Ei, F, V, status = FusedDP(w, "bf16", dev)((nl.frac_wrapped @ L).contiguous(), L.reshape(1, 9).contiguous(), tm,
                                               nl.list, nl.shift, nl.counts, w.cfg.m, rev=rev)
    torch.cuda.synchronize()
    assert int(status.item()) == 0 and int(rst.item()) == 0
    rmse = float(np.sqrt(((F.double().cpu().numpy() - ref["F"]) ** 2).mean()))
    assert rmse <= 1e-3, rmse
