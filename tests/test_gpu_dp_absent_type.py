"""Frames whose atom_type list lacks a model type, packed with frames that have it.

The reference forms the (ti, tj) embedding pairs only over the types of each image's atom_type
list and divides the descriptor by m times their number (src/model/dp_dp.py:150-165, 197, 413);
oracle.dp_evaluate(type_present=...) restates that, and tests/test_oracle_dp.py pins it against a
model with the absent type removed. Here the HIP paths are held to it, frame by frame:
  * fused E/F/V path: FP32 within 1e-6 |E| and 1e-5 eV/A per force component, BF16 force RMSE
    <= 1e-3 eV/A, against the FP64 oracle of each frame with its own type list;
  * plain FP64 path (DPTorch with a per-atom type list) on the packed batch: loss equal to the
    oracle's, and parameter gradients equal to central differences of the oracle's loss;
  * batched training step (DPTrain): loss and every parameter gradient within the precision of
    each mode of the plain FP64 path's on the same batch.
"""
import copy

import numpy as np
import pytest
import torch

import oracle
from matpl_amd import structures as st
from matpl_amd.dp import DPTorch
from matpl_amd.dp_fused import FusedDP, frame_types_of
from matpl_amd.dp_train import DPTrain, pack_batch, plain_grads, train_grads
from matpl_amd.dpnet import DPConfig, DPWeights
from matpl_amd.neighbor import find_neighbore_batch, reverse_map

pytestmark = pytest.mark.gpu

ELEM = ["Li", "P", "S"]


def _frames(seed=5):
    """Three-type frames: all types, type 1 absent, type 0 absent, all types again."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = [[40, 20, 36], [50, 0, 42], [0, 56, 40], [30, 30, 30]]
    return [st._random_frame(sum(c), c, ELEM, 0.06, 1.9, rng) for c in counts]


def _weights(frames, seed=11):
    cfg = DPConfig(ntypes=3, m=100)
    f = frames[0]
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 3, cfg.m, cfg.rcut)
    davg, dstd = oracle.dp_calc_stat(lst, dR, f.types, 3, cfg.rcut_smooth, cfg.rcut)
    return DPWeights.random(cfg, seed=seed, davg=davg, dstd=dstd)


def _occurring(f):
    tp = np.zeros(3, np.int32)
    tp[np.unique(f.types)] = 1
    return tp


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_fused_mixed_type_lists(dev, precision):
    frames = _frames()
    w = _weights(frames)
    m = w.cfg.m
    print(f"case fused absent-type {precision}: sizes {[f.natoms for f in frames]}", flush=True)
    nl = find_neighbore_batch([f.lattice for f in frames], [f.frac() for f in frames], [f.types for f in frames],
                              3, m, w.cfg.rcut, want_dR=False, want_shift=True, device=dev)
    sizes = [f.natoms for f in frames]
    fs = torch.as_tensor(np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64), device=dev)
    af = torch.as_tensor(np.repeat(np.arange(len(frames), dtype=np.int32), sizes), device=dev)
    lat = torch.as_tensor(np.stack([np.asarray(f.lattice).reshape(9) for f in frames]), device=dev)
    pos = torch.cat([nl.frac_wrapped[int(fs[k]):int(fs[k + 1])] @ lat[k].reshape(3, 3)
                     for k in range(len(frames))]).contiguous()
    tm = torch.as_tensor(np.concatenate([f.types for f in frames]), device=dev)
    rev, rst = reverse_map(nl, tm, m, frame_start=fs, atom_frame=af)
    assert int(rst.item()) == 0
    tp = frame_types_of(tm, af, len(frames), 3)
    assert not bool(tp.all())                           # some frame lacks a type
    model = FusedDP(w, precision, dev)
    Ei, F, V, status = model(pos, lat, tm, nl.list, nl.shift, nl.counts, m, atom_frame=af, frame_start=fs,
                             nframes=len(frames), rev=rev, type_present=tp)
    torch.cuda.synchronize()
    assert int(status.item()) == 0
    Ei, F, V = Ei.double().cpu().numpy(), F.double().cpu().numpy(), V.cpu().numpy()
    for k, f in enumerate(frames):
        a, b = int(fs[k]), int(fs[k + 1])
        ref = oracle.dp_evaluate(f, w, type_present=_occurring(f))
        if precision == "fp32":
            assert abs(Ei[a:b].sum() - ref["E"]) <= 1e-6 * abs(ref["E"]), k
            assert np.abs(F[a:b] - ref["F"]).max() <= 1e-5, (k, np.abs(F[a:b] - ref["F"]).max())
            np.testing.assert_allclose(V[k].reshape(3, 3).T, ref["V"].reshape(3, 3),
                                       atol=1e-5 * np.abs(ref["V"]).max())
        else:
            rmse = np.sqrt(((F[a:b] - ref["F"]) ** 2).mean())
            assert rmse <= 1e-3, (k, rmse)
            assert abs(Ei[a:b].sum() - ref["E"]) <= 1e-3 * abs(ref["E"]), k


def test_type_list_must_cover_the_frames_atoms(dev):
    frames = _frames()
    w = _weights(frames)
    f = frames[0]
    nl = find_neighbore_batch([f.lattice], [f.frac()], [f.types], 3, w.cfg.m, w.cfg.rcut, want_dR=False,
                              want_shift=True, device=dev)
    L = torch.as_tensor(f.lattice, device=dev)
    tm = torch.as_tensor(f.types, device=dev)
    model = FusedDP(w, "fp32", dev)
    with pytest.raises(ValueError):
        model((nl.frac_wrapped @ L).contiguous(), L.reshape(1, 9).contiguous(), tm, nl.list, nl.shift, nl.counts,
              w.cfg.m, type_present=[1, 0, 1])
    with pytest.raises(ValueError):
        model((nl.frac_wrapped @ L).contiguous(), L.reshape(1, 9).contiguous(), tm, nl.list, nl.shift, nl.counts,
              w.cfg.m, type_present=[0, 0, 0])


def _oracle_loss(frames, w, batch):
    """loss_fn of matpl_amd.dp_train on the FP64 oracle's per-frame energies and forces."""
    E_lab = batch.E_label.cpu().numpy()
    F_lab = batch.F_label.cpu().numpy()
    le, lf, off = 0.0, 0.0, 0
    for k, f in enumerate(frames):
        r = oracle.dp_evaluate(f, w, type_present=_occurring(f))
        le += ((r["E"] - E_lab[k]) / f.natoms) ** 2
        lf += ((r["F"] - F_lab[off:off + f.natoms]) ** 2).sum()
        off += f.natoms
    return le / len(frames) + lf / (3 * off)


def test_plain_path_gradients_vs_oracle_finite_differences(dev):
    frames = _frames()
    w = _weights(frames)
    batch = pack_batch(frames, w.cfg, dev, seed=3, keep_full=True)
    assert not bool(batch.type_present.all())
    print(f"case plain absent-type gradients: N={batch.natoms}", flush=True)
    model = DPTorch(w, torch.float64, dev)
    loss, grads = plain_grads(model, batch)
    params = [p for p in model.parameters() if p.requires_grad]
    assert abs(loss.item() - _oracle_loss(frames, w, batch)) <= 1e-10 * abs(loss.item())
    named = dict(model.named_parameters())
    gmap = {n: g for (n, _), g in zip(model.named_parameters(), grads)}
    assert len(params) == len(grads)
    # (parameter name, numpy location in w, index): embedding pairs with and without an absent type
    picks = [("fit.0.W.1", ("fit", 0, 1, 0), (3, 7)), ("fit.2.b.0", ("fit", 2, 0, 1), (5,)),
             ("emb.1.W.1", ("emb", 1, 1, 0), (2, 4)), ("emb.5.b.2", ("emb", 5, 2, 1), (9,)),
             ("emb.6.W.0", ("emb", 6, 0, 0), (0, 11)), ("emb.4.W.2", ("emb", 4, 2, 0), (7, 1))]
    h = 1e-5
    for name, (kind, k, l, which), idx in picks:
        assert name in named, name
        vals = []
        for sgn in (1, -1):
            w2 = copy.deepcopy(w)
            nets = w2.fit if kind == "fit" else w2.emb
            layer = list(nets[k][l])
            arr = np.array(layer[which], copy=True)
            arr[idx] += sgn * h
            layer[which] = arr
            nets[k][l] = tuple(layer)
            vals.append(_oracle_loss(frames, w2, batch))
        fd = (vals[0] - vals[1]) / (2 * h)
        g = gmap[name]
        an = 0.0 if g is None else float(g[idx].item())
        assert abs(an - fd) <= 1e-6 * max(1e-3, abs(fd)), (name, idx, an, fd)


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-4), ("bf16", 3e-2)])
def test_batched_step_mixed_type_lists(dev, precision, tol):
    frames = _frames()
    w = _weights(frames)
    batch = pack_batch(frames, w.cfg, dev, seed=3, keep_full=True)
    print(f"case batched absent-type {precision}: N={batch.natoms} mc={batch.mc}", flush=True)
    ref_loss, ref_g = plain_grads(DPTorch(w, torch.float64, dev, tall=True), batch)
    loss, g = train_grads(DPTrain(w, precision, dev), batch)
    assert abs(loss.item() - ref_loss.item()) < tol * abs(ref_loss.item())
    worst = 0.0
    for a, b in zip(g, ref_g):
        assert (a is None) == (b is None)
        if a is not None and b.norm() > 0:
            worst = max(worst, float((a.double() - b.double()).norm() / b.double().norm()))
    print(f"worst parameter-gradient relative error {worst:.2e}", flush=True)
    assert worst < tol
