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
