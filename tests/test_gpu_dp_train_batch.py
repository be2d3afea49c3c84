"""The batched training step (matpl_amd.dp_train.DPTrain: packed frames, compact buckets with the
closed-form empty-slot term, FP32 descriptors, FP32 or BF16 products) against the plain FP64 path
(DPTorch over all T*m slots, whose force-loss gradients test_gpu_dp_train.py pins against finite
differences of the oracle): same loss, same parameter gradients, within the precision of each mode."""
import numpy as np
import pytest
import torch

from matpl_amd import structures as st
from matpl_amd.dp import DPTorch
from matpl_amd.dp_train import DPTrain, pack_batch, plain_grads, train_grads
from matpl_amd.dpnet import DPConfig, DPWeights
from matpl_amd.neighbor import calc_stat

pytestmark = pytest.mark.gpu


def _setup(dev, nframes=5, seed=21):
    frames = st.lgps_frames(nframes=nframes, nmin=24, nmax=72, seed=seed)
    cfg = DPConfig(ntypes=4, m=150)
    batch = pack_batch(frames, cfg, dev, seed=seed, keep_full=True)
    lst = batch.list_full[0]
    dR = batch.image_dr_full[0, ..., 1:]
    davg, dstd = calc_stat(lst, dR, batch.type_map, 4, cfg.rcut_smooth, cfg.rcut)
    w = DPWeights.random(cfg, seed=seed + 1, davg=davg.cpu().numpy(), dstd=dstd.cpu().numpy())
    return frames, cfg, batch, w


def _rel(a, b):
    return float((a.double() - b.double()).norm() / max(b.double().norm().item(), 1e-30))


def test_packed_plain_path_equals_per_frame(dev):
    frames, cfg, batch, w = _setup(dev, nframes=3)
    print(f"case packed vs per-frame: N={batch.natoms} mc={batch.mc}", flush=True)
    model = DPTorch(w, torch.float64, dev)
    l_all, g_all = plain_grads(model, batch)
    l_tall, g_tall = plain_grads(DPTorch(w, torch.float64, dev, tall=True), batch)
    assert abs(l_all.item() - l_tall.item()) < 1e-12 * abs(l_all.item())
    for a, b in zip(g_all, g_tall):
        if a is not None:
            assert _rel(a, b) < 1e-10
    l_one, g_one = plain_grads(model, batch, chunk_frames=1)
    assert abs(l_all.item() - l_one.item()) < 1e-12 * abs(l_all.item())
    for a, b in zip(g_all, g_one):
        if a is not None:
            assert _rel(a, b) < 1e-10


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-4), ("bf16", 3e-2)])
def test_batched_step_matches_plain_fp64(dev, precision, tol):
    frames, cfg, batch, w = _setup(dev)
    print(f"case {precision}: N={batch.natoms} mc={batch.mc} (m={cfg.m})", flush=True)
    assert batch.mc < cfg.m          # the closed-form empty-slot term is exercised
    ref_loss, ref_g = plain_grads(DPTorch(w, torch.float64, dev, tall=True), batch)
    loss, g = train_grads(DPTrain(w, precision, dev), batch)
    assert abs(loss.item() - ref_loss.item()) < tol * abs(ref_loss.item())
    worst = 0.0
    for a, b in zip(g, ref_g):
        assert (a is None) == (b is None)
        if a is not None and b.norm() > 0:
            worst = max(worst, _rel(a, b))
    print(f"worst parameter-gradient relative error {worst:.2e}", flush=True)
    assert worst < tol
