"""Host-side checks of the training step's per-step weight packing (no GPU)."""

def test_training_weight_packing_equals_host_packing():
    """The per-step device-side packing of the training step (matpl_amd.dp_train_fused) builds the
    fused kernels' operand layouts exactly as the host packing (matpl_amd.dp_fused.pack_fused)."""
    import numpy as np
    import torch
    from matpl_amd.dp_fused import pack_fused
    from matpl_amd.dp_train_fused import DPTrainFused
    from matpl_amd.dpnet import DPConfig, DPWeights
    rng = np.random.default_rng(1)
    for T in (1, 2, 4):
        cfg = DPConfig(ntypes=T, m=100)
        w = DPWeights.random(cfg, seed=3, davg=rng.uniform(0.1, 0.2, (T, 4)), dstd=rng.uniform(0.5, 1.0, (T, 4)))
        for prec in ("fp32", "bf16"):
            ref = pack_fused(w, prec, "cpu")
            vec, mat = DPTrainFused(w, prec, "cpu")._pack()
            assert torch.equal(mat.float(), ref.mat.float())
            assert torch.allclose(vec, ref.vec, rtol=0, atol=2e-6)


def test_fitting_weight_packing_equals_host_packing():
    """The training step's device-side packing of the fitting nets (one gather from the concatenated
    parameters) builds the matrix-core fitting kernels' operands exactly as the host packing
    (matpl_amd.dp_fitting.FittingMFMA)."""
    import numpy as np
    import torch
    from matpl_amd.dp import _Net
    from matpl_amd.dp_fitting import FittingMFMA
    from matpl_amd.dp_train_fused import _FitPack
    from matpl_amd.dpnet import DPConfig, DPWeights
    rng = np.random.default_rng(1)
    for T in (1, 2, 4):
        cfg = DPConfig(ntypes=T, m=100)
        w = DPWeights.random(cfg, seed=3, davg=rng.uniform(0.1, 0.2, (T, 4)), dstd=rng.uniform(0.5, 1.0, (T, 4)))
        nets = [_Net(w.fit[t], 1, torch.float32, "cpu") for t in range(T)]
        H, L = w.fit[0][0][0].shape[1], len(w.fit[0]) - 1
        for prec in ("fp32", "bf16"):
            ref = FittingMFMA(w, [(0, 0, 1)], prec, "cpu")
            w0f, w0b, whf, whb, vec, strides = _FitPack(cfg.emb_out, H, L, T, nets[0].has_rdt, prec == "bf16",
                                                        "cpu")(nets)
            for a, b in ((w0f, ref.w0f), (w0b, ref.w0b), (whf, ref.whf), (whb, ref.whb)):
                assert a.shape == b.shape and torch.equal(a.float(), b.float())
            assert torch.allclose(vec, ref.vec, rtol=0, atol=1e-7)
            assert tuple(strides) == tuple(ref.strides)


def test_batch_constants_follow_ri_pad():
    """The step's per-batch constants (kept on the batch) are reused while the model's Ri_pad is
    unchanged and renewed after load_state_dict changes it in place."""
    from types import SimpleNamespace
    import numpy as np
    import torch
    from matpl_amd.dp_train_fused import DPTrainFused
    from matpl_amd.dpnet import DPConfig, DPWeights
    cfg = DPConfig(ntypes=2, m=60)
    w = DPWeights.random(cfg, seed=3, davg=np.tile([0.1, 0, 0, 0], (2, 1)), dstd=np.tile([0.2, 0.1, 0.1, 0.1], (2, 1)))
    model = DPTrainFused(w, "fp32", "cpu")
    model._fit_tiles = lambda b: (torch.arange(3, dtype=torch.int32), None, None)
    b = SimpleNamespace(atom_frame=torch.zeros(3, dtype=torch.int32), type_present=torch.ones(1, 2, dtype=torch.bool),
                        type_map=torch.tensor([0, 1, 1], dtype=torch.int32))
    k1 = model._batch_const(b)
    assert model._batch_const(b) is k1
    assert torch.equal(k1["rp"][:, :, 0], model.ri_pad[b.type_map.long()])
    sd = model.state_dict()
    sd["ri_pad"] = sd["ri_pad"] * 2
    model.load_state_dict(sd)
    k2 = model._batch_const(b)
    assert k2 is not k1 and torch.equal(k2["rp"], 2 * k1["rp"])
