"""DP training step (force-loss double backward) on the plain drop-in path.

The loss L = sum F^2 + E^2 + sum V^2 of DPTorch (torch nets + the HIP
CalcOps_cuda force/virial ops and their backward) is differentiated to the
embedding and fitting parameters. Each gradient entry is checked against a
central finite difference of the same loss evaluated by the FP64 oracle
(forces and virial from oracle_dp_force_virial), so the op backward, the
env-mat derivative and the double backward through the nets are all pinned
against an independent forward computation.
"""
import copy

import numpy as np
import pytest
import torch

import oracle
from matpl_amd import structures as st
from matpl_amd.dpnet import DPConfig, DPWeights

pytestmark = pytest.mark.gpu


def _loss_oracle(f, w):
    r = oracle.dp_evaluate(f, w)
    V = r["V"].reshape(3, 3).T     # GPU op convention (transpose of the CPU branch)
    return float((r["F"] ** 2).sum() + r["E"] ** 2 + (V ** 2).sum())


def test_force_loss_parameter_gradients_vs_finite_difference(dev):
    from matpl_amd.dp import DPTorch
    from matpl_amd.neighbor import find_neighbore_gpu, image_dr
    f = st.fluorite_hfo2((2, 2, 2), seed=31)
    cfg = DPConfig(ntypes=2, m=100)
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 2, cfg.m, cfg.rcut)
    davg, dstd = oracle.dp_calc_stat(lst, dR, f.types, 2, cfg.rcut_smooth, cfg.rcut)
    w = DPWeights.random(cfg, seed=6, davg=davg, dstd=dstd)
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, cfg.m, cfg.rcut, device=dev)
    model = DPTorch(w, torch.float64, dev)
    Etot, Ei, F, V = model(nl.list.unsqueeze(0), torch.as_tensor(f.types, device=dev),
                           image_dr(nl.list, nl.dR).unsqueeze(0), create_graph=True)
    loss = (F ** 2).sum() + (Etot ** 2).sum() + (V ** 2).sum()
    assert abs(loss.item() - _loss_oracle(f, w)) < 1e-10 * abs(loss.item())
    # (net list, net index, layer, which, index) entries to check
    picks = [("fit", 0, 0, 0, (17, 3)), ("fit", 1, 1, 0, (4, 9)), ("fit", 0, 2, 1, (5,)), ("fit", 1, 3, 0, (7, 0)),
             ("fit", 0, 1, 2, (11,)), ("emb", 0, 0, 0, (0, 3)), ("emb", 3, 1, 0, (2, 20)), ("emb", 1, 2, 1, (8,)),
             ("emb", 2, 0, 1, (24,))]
    params = []
    for kind, k, l, which, idx in picks:
        net = (model.fit if kind == "fit" else model.emb)[k]
        params.append((net.W, net.b, net.r)[which][l])
    grads = torch.autograd.grad(loss, params)
    h = 1e-5
    for (kind, k, l, which, idx), g in zip(picks, grads):
        vals = []
        for sgn in (1, -1):
            w2 = copy.deepcopy(w)
            nets = w2.fit if kind == "fit" else w2.emb
            layer = list(nets[k][l])
            arr = np.array(layer[which], copy=True)
            arr[idx] += sgn * h
            layer[which] = arr
            nets[k][l] = tuple(layer)
            vals.append(_loss_oracle(f, w2))
        fd = (vals[0] - vals[1]) / (2 * h)
        an = float(g[idx].item())
        assert abs(an - fd) <= 1e-6 * max(1.0, abs(fd)), (kind, k, l, which, idx, an, fd)
