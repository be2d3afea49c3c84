"""The batched DP training step on the hand-written kernels (matpl_amd.dp_train_fused.DPTrainFused:
fused forward and force kernels, the force-loss double backward of the embedding and fitting nets on
the matrix cores; the fitting side also on torch autograd) against the plain FP64 path (DPTorch over
all T*m slots with create_graph, matpl_amd.dp_train.plain_grads), whose gradients the last test pins
against central differences of the FP64 oracle's loss:
  loss within 1e-5 (FP32) / 3e-3 (BF16) relative, every parameter gradient within 2e-4 (FP32) /
  3e-2 (BF16) relative norm; frames of Li-Ge-P-S with N in [24, 512], and frames whose type lists
  differ (one lacks a type)."""
import copy

import numpy as np
import pytest
import torch

import oracle
from matpl_amd import structures as st
from matpl_amd.dp import DPTorch
from matpl_amd.dp_train import pack_batch, plain_grads
from matpl_amd.dp_train_fused import DPTrainFused, pack_fused_batch
from matpl_amd.dpnet import DPConfig, DPWeights
from matpl_amd.neighbor import calc_stat

pytestmark = pytest.mark.gpu

TOL = {"fp32": (1e-5, 2e-4), "bf16": (3e-3, 3e-2)}


def _weights(frames, cfg, dev, seed):
    b = pack_batch(frames, cfg, dev, seed=0, keep_full=True)
    davg, dstd = calc_stat(b.list_full[0], b.image_dr_full[0, ..., 1:], b.type_map, cfg.ntypes, cfg.rcut_smooth,
                           cfg.rcut)
    return DPWeights.random(cfg, seed=seed, davg=davg.cpu().numpy(), dstd=dstd.cpu().numpy())


def _compare(frames, cfg, w, dev, precision, seed=3, fitting="kernel"):
    ref_batch = pack_batch(frames, cfg, dev, seed=seed, keep_full=True)
    ref_loss, ref_g = plain_grads(DPTorch(w, torch.float64, dev), ref_batch)
    fb = pack_fused_batch(frames, cfg, dev, seed=seed)
    assert torch.equal(fb.E_label, ref_batch.E_label) and torch.equal(fb.F_label, ref_batch.F_label)
    model = DPTrainFused(w, precision, dev, fitting=fitting)
    loss, g = model.loss_and_grads(fb)
    tl, tg = TOL[precision]
    print(f"loss {loss.item():.8e} plain FP64 {ref_loss.item():.8e}", flush=True)
    assert abs(loss.item() - ref_loss.item()) <= tl * abs(ref_loss.item())
    names = [n for n, _ in model.named_parameters()]
    assert len(g) == len(ref_g) == len(names)
    worst, where = 0.0, None
    params = [p for _, p in model.named_parameters()]
    for name, a, b, prm in zip(names, g, ref_g, params):
        assert (a is None) == (b is None), name
        # laid out like the parameter, so that the optimizer's multi-tensor kernels apply
        assert a is None or (a.shape == prm.shape and a.stride() == prm.stride()), name
        if a is not None and b.norm() > 0:
            e = float((a.double() - b.double()).norm() / b.double().norm())
            if e > worst:
                worst, where = e, name
    print(f"{precision}: worst parameter-gradient relative error {worst:.2e} ({where})", flush=True)
    assert worst <= tg, where


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_fused_step_matches_plain_fp64(dev, precision):
    frames = st.lgps_frames(nframes=5, nmin=24, nmax=72, seed=21)
    cfg = DPConfig(ntypes=4, m=150)
    w = _weights(frames, cfg, dev, seed=22)
    print(f"case fused training step {precision}: sizes {[f.natoms for f in frames]}", flush=True)
    _compare(frames, cfg, w, dev, precision)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_fused_step_torch_fitting_matches_plain_fp64(dev, precision):
    """The same step with the fitting nets' derivatives from torch autograd instead of the kernels."""
    frames = st.lgps_frames(nframes=5, nmin=24, nmax=72, seed=21)
    cfg = DPConfig(ntypes=4, m=150)
    w = _weights(frames, cfg, dev, seed=22)
    print(f"case fused training step, torch fitting {precision}", flush=True)
    _compare(frames, cfg, w, dev, precision, fitting="torch")


def test_fitting_kernels_match_torch_fitting(dev):
    """FP32: the fitting kernels' loss and gradients against the torch-fitting path of the same step
    (both FP32, same embedding kernels): loss within 1e-6, every gradient within 1e-4 relative norm."""
    frames = st.lgps_frames(nframes=4, nmin=64, nmax=200, seed=8)
    cfg = DPConfig(ntypes=4, m=150)
    w = _weights(frames, cfg, dev, seed=9)
    fb = pack_fused_batch(frames, cfg, dev, seed=3)
    print("case fitting kernels vs torch fitting (fp32)", flush=True)
    l1, g1 = DPTrainFused(w, "fp32", dev, fitting="kernel").loss_and_grads(fb)
    l2, g2 = DPTrainFused(w, "fp32", dev, fitting="torch").loss_and_grads(fb)
    assert abs(l1.item() - l2.item()) <= 1e-6 * abs(l2.item())
    names = [n for n, _ in DPTrainFused(w, "fp32", dev).named_parameters()]
    worst = 0.0
    for name, a, b in zip(names, g1, g2):
        assert (a is None) == (b is None), name
        if a is not None and b.norm() > 0:
            e = float((a - b).norm() / b.norm())
            worst = max(worst, e)
            assert e <= 1e-4, (name, e)
    print(f"worst relative difference {worst:.2e}", flush=True)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_fused_step_config_sizes(dev, precision):
    """Frames in the configuration's size range [64, 512]."""
    frames = st.lgps_frames(nframes=3, nmin=64, nmax=512, seed=1002)
    cfg = DPConfig(ntypes=4, m=150)
    w = _weights(frames, cfg, dev, seed=1002)
    print(f"case fused training step {precision}: sizes {[f.natoms for f in frames]}", flush=True)
    _compare(frames, cfg, w, dev, precision)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_fused_step_mixed_type_lists(dev, precision):
    rng = np.random.Generator(np.random.PCG64(5))
    frames = [st._random_frame(sum(c), c, ["Li", "P", "S"], 0.06, 1.9, rng)
              for c in ([40, 20, 36], [50, 0, 42], [0, 56, 40], [30, 30, 30])]
    cfg = DPConfig(ntypes=3, m=100)
    w = _weights(frames, cfg, dev, seed=11)
    print(f"case fused training step, mixed type lists {precision}", flush=True)
    _compare(frames, cfg, w, dev, precision)


def test_plain_path_vs_oracle_finite_differences_lgps(dev):
    """The reference for the two tests above: the plain FP64 path's loss and gradients on Li-Ge-P-S
    frames of the configuration's sizes against the FP64 oracle's loss and its central differences
    (h = 1e-5 in the parameters)."""
    frames = st.lgps_frames(nframes=2, nmin=64, nmax=512, seed=1002)
    cfg = DPConfig(ntypes=4, m=150)
    w = _weights(frames, cfg, dev, seed=1002)
    batch = pack_batch(frames, cfg, dev, seed=3, keep_full=True)
    print(f"case plain-path gradients vs oracle differences: sizes {[f.natoms for f in frames]}", flush=True)
    model = DPTorch(w, torch.float64, dev)
    loss, grads = plain_grads(model, batch)
    E_lab, F_lab = batch.E_label.cpu().numpy(), batch.F_label.cpu().numpy()

    def oracle_loss(ww):
        le, lf, off = 0.0, 0.0, 0
        for k, f in enumerate(frames):
            tp = np.zeros(cfg.ntypes, np.int32)
            tp[np.unique(f.types)] = 1
            r = oracle.dp_evaluate(f, ww, type_present=tp)
            le += ((r["E"] - E_lab[k]) / f.natoms) ** 2
            lf += ((r["F"] - F_lab[off:off + f.natoms]) ** 2).sum()
            off += f.natoms
        return le / len(frames) + lf / (3 * off)

    assert abs(loss.item() - oracle_loss(w)) <= 1e-10 * abs(loss.item())
    gmap = {n: g for (n, _), g in zip(model.named_parameters(), grads)}
    picks = [("fit.1.W.1", ("fit", 1, 1, 0), (3, 7)), ("fit.3.b.2", ("fit", 3, 2, 1), (5,)),
             ("emb.5.W.1", ("emb", 5, 1, 0), (2, 4)), ("emb.14.b.2", ("emb", 14, 2, 1), (9,)),
             ("emb.0.W.0", ("emb", 0, 0, 0), (0, 11))]
    h = 1e-5
    for name, (kind, k, l, which), idx in picks:
        vals = []
        for sgn in (1, -1):
            w2 = copy.deepcopy(w)
            nets = w2.fit if kind == "fit" else w2.emb
            layer = list(nets[k][l])
            arr = np.array(layer[which], copy=True)
            arr[idx] += sgn * h
            layer[which] = arr
            nets[k][l] = tuple(layer)
            vals.append(oracle_loss(w2))
        fd = (vals[0] - vals[1]) / (2 * h)
        an = float(gmap[name][idx].item())
        print(f"{name}{idx}: autograd {an:.8e} central difference {fd:.8e}", flush=True)
        assert abs(an - fd) <= 1e-6 * max(1e-3, abs(fd)), (name, idx, an, fd)
