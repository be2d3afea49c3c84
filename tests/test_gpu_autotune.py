"""Launch shapes of the fused DP kernels (matpl_dp_fused_set_launch, matpl_dp_fused_set_waves) change
speed, not results. Within one instantiation (wavefronts per block) every grid cap, weight placement
(LDS or global) and block order gives the same energies and forces bit for bit (gather mode sums in a
fixed order) and the virial to FP64 summation order. The 2, 4 and 8 wavefront instantiations are
compiled separately and may round differently: their forces and virials agree within 1e-6 of the
largest component (FP32 rounding: measured 3.7e-9 against a largest force component of 0.0315,
1.2e-7, and 2.1e-8 on the virial, at 8 waves), energies bit for bit. The autotuner picks
a shape, sets it and caches it."""
import numpy as np
import pytest
import torch

from matpl_amd import autotune
from matpl_amd import structures as st
from matpl_amd.dp_fused import FusedDP
from matpl_amd.dpnet import DPConfig, DPWeights
from matpl_amd.neighbor import find_neighbore_gpu, reverse_map

pytestmark = pytest.mark.gpu


def _workload(dev, precision="bf16"):
    f = st.fluorite_hfo2((3, 3, 3), seed=21)
    cfg = DPConfig(ntypes=2, m=120)
    w = DPWeights.random(cfg, seed=5, davg=np.tile([0.1, 0, 0, 0], (2, 1)), dstd=np.tile([0.2, 0.1, 0.1, 0.1], (2, 1)))
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, cfg.m, cfg.rcut, device=dev, want_dR=False,
                            want_shift=True)
    L = torch.as_tensor(f.lattice, device=dev)
    pos = (nl.frac_wrapped @ L).contiguous()
    tm = torch.as_tensor(f.types, device=dev)
    rev, rst = reverse_map(nl, tm, cfg.m)
    assert int(rst.item()) == 0
    model = FusedDP(w, precision, dev)
    ranges = [(int(np.searchsorted(f.types, t, "left")), int(np.searchsorted(f.types, t, "right"))) for t in range(2)]

    def run():
        return model(pos, L.reshape(1, 9).contiguous(), tm, nl.list, nl.shift, nl.counts, cfg.m,
                     type_ranges=ranges, rev=rev)
    return f, cfg, run


def test_results_independent_of_launch_shape(dev):
    f, cfg, run = _workload(dev)
    print(f"case launch shapes: N={f.natoms}", flush=True)
    try:
        autotune.reset()
        Ei0, F0, V0, st0 = [x.clone() for x in run()]
        for k in (0, 1):
            for cand in ({"max_blocks": 8, "lds_weights": 0, "xcd": 1}, {"max_blocks": 3, "lds_weights": -1, "xcd": 1},
                         {"max_blocks": 64, "lds_weights": 1, "xcd": 0}, {"max_blocks": 16384, "lds_weights": 0,
                                                                          "xcd": 1},
                         {"max_blocks": 5, "lds_weights": 1, "xcd": 0, "waves": 2},
                         {"max_blocks": 4096, "lds_weights": -1, "xcd": 1, "waves": 8},
                         {"max_blocks": 7, "lds_weights": 0, "xcd": 0, "waves": 8}):
                autotune.set_launch(k, **cand)
                Ei, F, V, status = run()
                torch.cuda.synchronize()
                assert int(status.item()) == 0
                assert torch.equal(Ei, Ei0), (k, cand)
                if cand.get("waves", 4) == 4:
                    assert torch.equal(F, F0), (k, cand)
                    assert torch.allclose(V, V0, rtol=1e-12, atol=1e-12 * float(V0.abs().max())), (k, cand)
                else:
                    assert float((F - F0).abs().max()) <= 1e-6 * float(F0.abs().max()), (k, cand)
                    assert float((V - V0).abs().max()) <= 1e-6 * float(V0.abs().max()), (k, cand)
            autotune.set_launch(k, **autotune.DEFAULT)
    finally:
        autotune.reset()


def test_autotuner_sets_and_caches(dev, tmp_path, monkeypatch):
    monkeypatch.setenv("MATPL_TUNE_CACHE", str(tmp_path / "tune.json"))
    f, cfg, run = _workload(dev)
    print("case autotune", flush=True)
    try:
        key = autotune.workload_key(25, 2, cfg.m, cfg.m, f.natoms, "bf16")
        r1 = autotune.tune_fused(run, key, reps=2)
        assert not r1["cached"]
        # timings of two repetitions on a shared GPU are noisy: logged, not asserted
        print(f"tuned {r1['ms']:.3f} ms against default {r1['ms_default']:.3f} ms", flush=True)
        assert (tmp_path / "tune.json").exists()
        r2 = autotune.tune_fused(run, key, reps=2)
        assert r2["cached"] and r2["forward"] == r1["forward"] and r2["backward"] == r1["backward"]
    finally:
        autotune.reset()


def test_waves_fall_back_when_tiles_exceed_lds(dev):
    """A wavefront count set for one workload (8 waves, tuned on a two-type box) does not break
    another whose neighbor tiles would not fit the LDS budget at that count (four types, m = 150):
    the launch falls back to fewer wavefronts, and the result equals the default launch's within
    FP32 rounding."""
    f = st.lgps_frames(nframes=1, nmin=200, nmax=200, seed=3)[0]
    T, m = 4, 150
    cfg = DPConfig(ntypes=T, m=m)
    w = DPWeights.random(cfg, seed=5, davg=np.tile([0.1, 0, 0, 0], (T, 1)), dstd=np.tile([0.2, 0.1, 0.1, 0.1], (T, 1)))
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, T, m, cfg.rcut, device=dev, want_dR=False, want_shift=True)
    L = torch.as_tensor(f.lattice, device=dev)
    pos = (nl.frac_wrapped @ L).contiguous()
    tm = torch.as_tensor(f.types, device=dev)
    rev, rst = reverse_map(nl, tm, m)
    assert int(rst.item()) == 0
    model = FusedDP(w, "bf16", dev)
    print(f"case waves fallback: N={f.natoms} T={T} m={m}", flush=True)

    def run():
        out = model(pos, L.reshape(1, 9).contiguous(), tm, nl.list, nl.shift, nl.counts, m, rev=rev)
        torch.cuda.synchronize()
        return [x.clone() for x in out]
    try:
        autotune.reset()
        Ei0, F0, V0, s0 = run()
        for k in (0, 1):
            autotune.set_launch(k, 4096, -1, 0, 8)
        Ei, F, V, s = run()
        assert int(s.item()) == 0
        assert float((Ei - Ei0).abs().max()) <= 1e-6 * float(Ei0.abs().max())
        assert float((F - F0).abs().max()) <= 1e-6 * float(F0.abs().max())
    finally:
        autotune.reset()
