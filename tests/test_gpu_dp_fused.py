"""GPU parity of the fused DP path (positions + neighbor list -> E, F, V) vs the FP64 oracle.

Tolerances (stated per precision mode):
  fp32: |E - E64| <= 1e-6 |E64|, max |F - F64| <= 1e-5 eV/A, virial rel 1e-5
  bf16: force RMSE vs FP64 <= 1e-3 eV/A, |E - E64| <= 1e-3 |E64|
"""
import numpy as np
import pytest
import torch

import oracle
from matpl_amd import structures as st
from matpl_amd.dp_fused import FusedDP
from matpl_amd.dpnet import DPConfig, DPWeights
from matpl_amd.neighbor import find_neighbore_gpu, reverse_map

pytestmark = pytest.mark.gpu


def _setup(frame, T, m, seed, E=25):
    cfg = DPConfig(ntypes=T, m=m, emb_sizes=[E, E, E])
    lst, dR = oracle.dp_find_neighbor(frame.lattice, frame.frac(), frame.types, T, m, cfg.rcut)
    davg, dstd = oracle.dp_calc_stat(lst, dR, frame.types, T, cfg.rcut_smooth, cfg.rcut)
    return DPWeights.random(cfg, seed=seed, davg=davg, dstd=dstd)


def _fused(frame, w, precision, dev, rc_nl=None, mnl=None, gather=True, store=False):
    cfg = w.cfg
    rc_nl = rc_nl or cfg.rcut
    mnl = mnl or cfg.m
    nl = find_neighbore_gpu(frame.lattice, frame.frac(), frame.types, cfg.ntypes, mnl, rc_nl, device=dev,
                            want_dR=False, want_shift=True)
    L = torch.as_tensor(frame.lattice, device=dev)
    pos = (nl.frac_wrapped @ L).contiguous()
    model = FusedDP(w, precision, dev, store=store)
    tm = torch.as_tensor(frame.types, device=dev)
    rev = None
    if gather:
        rev, rst = reverse_map(nl, tm, mnl)
        assert int(rst.item()) == 0
    Ei, F, V, status = model(pos, L.reshape(1, 9).contiguous(), tm, nl.list, nl.shift, nl.counts, mnl, rev=rev)
    torch.cuda.synchronize()
    assert int(status.item()) == 0
    return Ei.double().cpu().numpy(), F.double().cpu().numpy(), V[0].cpu().numpy().reshape(3, 3)


CASES = {
    "cu108": (lambda: st.fcc_cu(), 1, 100),
    "hfo2_324": (lambda: st.fluorite_hfo2((3, 3, 3), seed=21), 2, 120),
    "sic_512": (lambda: st.zincblende_sic((4, 4, 4), seed=4), 2, 100),
}


@pytest.mark.parametrize("name", list(CASES))
def test_fused_fp32_matches_fp64_oracle(dev, name):
    maker, T, m = CASES[name]
    f = maker()
    w = _setup(f, T, m, seed=5)
    ref = oracle.dp_evaluate(f, w)
    Ei, F, V = _fused(f, w, "fp32", dev)
    assert abs(Ei.sum() - ref["E"]) <= 1e-6 * abs(ref["E"])
    assert np.abs(F - ref["F"]).max() <= 1e-5, np.abs(F - ref["F"]).max()
    np.testing.assert_allclose(V.T, ref["V"].reshape(3, 3), atol=1e-5 * np.abs(ref["V"]).max())


@pytest.mark.parametrize("name", list(CASES))
def test_fused_bf16_force_rmse(dev, name):
    maker, T, m = CASES[name]
    f = maker()
    w = _setup(f, T, m, seed=5)
    ref = oracle.dp_evaluate(f, w)
    Ei, F, V = _fused(f, w, "bf16", dev)
    rmse = np.sqrt(((F - ref["F"]) ** 2).mean())
    print(f"{name}: bf16 force RMSE {rmse:.2e} (|F| rms {np.sqrt((ref['F'] ** 2).mean()):.3f}), "
          f"dE/E {abs(Ei.sum() - ref['E']) / abs(ref['E']):.2e}")
    assert rmse <= 1e-3
    assert abs(Ei.sum() - ref["E"]) <= 1e-3 * abs(ref["E"])


def test_fused_skin_list_gives_same_result(dev):
    """Neighbors between rcut and the list cutoff act as empty slots."""
    f = st.fluorite_hfo2((4, 4, 4), seed=8)
    w = _setup(f, 2, 120, seed=9)
    Ei0, F0, V0 = _fused(f, w, "fp32", dev)
    Ei1, F1, V1 = _fused(f, w, "fp32", dev, rc_nl=7.0, mnl=200)
    np.testing.assert_allclose(Ei1, Ei0, atol=1e-6)
    np.testing.assert_allclose(F1, F0, atol=1e-6)
    np.testing.assert_allclose(V1, V0, atol=1e-5 * np.abs(V0).max())


def test_fused_overflow_flag(dev):
    f = st.fluorite_hfo2((4, 4, 4), seed=8)
    w = _setup(f, 2, 120, seed=9)
    cfg_small = DPConfig(ntypes=2, m=40)
    w_small = DPWeights(cfg_small, w.emb, w.fit, w.davg, w.dstd)
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, 120, 6.0, device=dev, want_dR=False, want_shift=True)
    L = torch.as_tensor(f.lattice, device=dev)
    model = FusedDP(w_small, "fp32", dev)
    _, _, _, status = model((nl.frac_wrapped @ L).contiguous(), L.reshape(1, 9).contiguous(),
                            torch.as_tensor(f.types, device=dev), nl.list, nl.shift, nl.counts, 120)
    assert int(status.item()) != 0


@pytest.mark.parametrize("name", ["cu108", "hfo2_324"])
def test_reverse_map_is_involution(dev, name):
    maker, T, m = CASES[name]
    f = maker()
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, T, m, 6.0, device=dev, want_dR=True, want_shift=True)
    rev, st_ = reverse_map(nl, torch.as_tensor(f.types, device=dev), m)
    assert int(st_.item()) == 0
    r = rev.reshape(-1).cpu().numpy()
    used = np.nonzero(r >= 0)[0]
    assert len(used) == int((nl.list > 0).sum().item())
    np.testing.assert_array_equal(r[r[used]], used)                 # rev(rev(s)) = s
    dR = nl.dR.reshape(-1, 3).cpu().numpy()
    np.testing.assert_array_equal(dR[r[used]], -dR[used])             # partner has the opposite displacement


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_gather_equals_atomic_scatter(dev, precision):
    f = st.fluorite_hfo2((4, 4, 4), seed=8)
    w = _setup(f, 2, 120, seed=9)
    Ei0, F0, V0 = _fused(f, w, precision, dev, gather=False)
    Ei1, F1, V1 = _fused(f, w, precision, dev, gather=True)
    np.testing.assert_allclose(Ei1, Ei0, atol=0)
    np.testing.assert_allclose(F1, F0, atol=2e-6 * max(1.0, np.abs(F0).max()))
    np.testing.assert_allclose(V1, V0, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-6), ("bf16", 3e-3)])
def test_fitting_kernel_matches_torch_autograd(dev, precision, tol):
    """Matrix-core fitting kernel (Ei, dEtot/dA) vs the FP32 torch fitting net with autograd."""
    from matpl_amd.dp_fitting import FittingMFMA
    from matpl_amd.dp_fused import _Fitting
    cfg = DPConfig(ntypes=2, m=120)
    w = DPWeights.random(cfg, seed=31)
    rng = np.random.default_rng(4)
    N = 203                                              # ragged: tiles of 32 plus a partial tile per type
    ranges = [(0, 77), (77, N)]
    A = torch.as_tensor(rng.normal(size=(N, 4, 25)) * 0.3, dtype=torch.float32, device=dev)
    Ei_ref, dA_ref = _Fitting(w, dev)(A, None, ranges)
    Ei, dA = FittingMFMA(w, ranges, precision, dev)(A)
    torch.cuda.synchronize()
    assert torch.allclose(Ei, Ei_ref, atol=tol * float(Ei_ref.abs().max()), rtol=0)
    assert torch.allclose(dA, dA_ref, atol=tol * float(dA_ref.abs().max()), rtol=0)


WIDTH_CASES = {
    "hfo2_324_E16": (lambda: st.fluorite_hfo2((3, 3, 3), seed=21), 2, 120, 16),
    "hfo2_324_E32": (lambda: st.fluorite_hfo2((3, 3, 3), seed=21), 2, 120, 32),
    "lgps_300_T4": (lambda: st.lgps_frames(nframes=1, nmin=300, nmax=300, seed=3)[0], 4, 150, 25),
}


@pytest.mark.parametrize("name", list(WIDTH_CASES))
def test_fused_other_widths_and_types(dev, name):
    """The other compiled embedding widths (16, 32) and four atom types (Li-Ge-P-S, 16 type pairs,
    weights read from global memory: they exceed the LDS budget) against the FP64 oracle: FP32 within
    1e-6 |E| and 1e-5 eV/A per force component, BF16 within force RMSE 1e-3 eV/A."""
    maker, T, m, E = WIDTH_CASES[name]
    f = maker()
    w = _setup(f, T, m, seed=7, E=E)
    ref = oracle.dp_evaluate(f, w)
    print(f"case fused {name}", flush=True)
    Ei, F, V = _fused(f, w, "fp32", dev)
    assert abs(Ei.sum() - ref["E"]) <= 1e-6 * abs(ref["E"])
    assert np.abs(F - ref["F"]).max() <= 1e-5, np.abs(F - ref["F"]).max()
    np.testing.assert_allclose(V.T, ref["V"].reshape(3, 3), atol=1e-5 * np.abs(ref["V"]).max())
    Ei, F, V = _fused(f, w, "bf16", dev)
    rmse = np.sqrt(((F - ref["F"]) ** 2).mean())
    print(f"{name}: bf16 force RMSE {rmse:.2e}", flush=True)
    assert rmse <= 1e-3
    assert abs(Ei.sum() - ref["E"]) <= 1e-3 * abs(ref["E"])


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_fused_several_frames(dev, precision):
    """Frames packed along the atom axis (atom_frame / frame_start, a lattice per frame, list indices
    within the frame, boxes small enough that several images of a neighbor appear): per-frame energies,
    forces and virials against the FP64 oracle of each frame (one atom type, so the packed atoms stay
    grouped by type). FP32 within 1e-6 |E| and 1e-5 eV/A; BF16 within force RMSE 1e-3 eV/A."""
    from matpl_amd.neighbor import find_neighbore_batch
    frames = [st.fcc_cu((3, 3, 3), seed=1), st.fcc_cu((3, 3, 4), a=3.65, seed=2), st.fcc_cu((4, 3, 3), seed=3)]
    T, m = 1, 100
    w = _setup(frames[0], T, m, seed=9)
    print(f"case fused {len(frames)} frames {precision}", flush=True)
    nl = find_neighbore_batch([f.lattice for f in frames], [f.frac() for f in frames], [f.types for f in frames],
                              T, m, w.cfg.rcut, want_dR=False, want_shift=True, device=dev)
    sizes = [f.natoms for f in frames]
    fs = torch.as_tensor(np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64), device=dev)
    af = torch.as_tensor(np.repeat(np.arange(len(frames), dtype=np.int32), sizes), device=dev)
    lat = torch.as_tensor(np.stack([np.asarray(f.lattice).reshape(9) for f in frames]), device=dev)
    pos = torch.cat([nl.frac_wrapped[int(fs[k]):int(fs[k + 1])] @ lat[k].reshape(3, 3)
                     for k in range(len(frames))]).contiguous()
    tm = torch.zeros(sum(sizes), dtype=torch.int32, device=dev)
    rev, rst = reverse_map(nl, tm, m, frame_start=fs, atom_frame=af)
    assert int(rst.item()) == 0
    model = FusedDP(w, precision, dev)
    Ei, F, V, status = model(pos, lat, tm, nl.list, nl.shift, nl.counts, m, atom_frame=af, frame_start=fs,
                             nframes=len(frames), type_ranges=[(0, sum(sizes))], rev=rev)
    torch.cuda.synchronize()
    assert int(status.item()) == 0
    Ei, F, V = Ei.double().cpu().numpy(), F.double().cpu().numpy(), V.cpu().numpy()
    for k, f in enumerate(frames):
        a, b = int(fs[k]), int(fs[k + 1])
        ref = oracle.dp_evaluate(f, w)
        if precision == "fp32":
            assert abs(Ei[a:b].sum() - ref["E"]) <= 1e-6 * abs(ref["E"]), k
            assert np.abs(F[a:b] - ref["F"]).max() <= 1e-5, (k, np.abs(F[a:b] - ref["F"]).max())
            np.testing.assert_allclose(V[k].reshape(3, 3).T, ref["V"].reshape(3, 3),
                                       atol=1e-5 * np.abs(ref["V"]).max())
        else:
            assert np.sqrt(((F[a:b] - ref["F"]) ** 2).mean()) <= 1e-3, k
            assert abs(Ei[a:b].sum() - ref["E"]) <= 1e-3 * abs(ref["E"]), k


STORE_CASES = {
    "cu108": (lambda: st.fcc_cu(), 1, 100, 25),
    "hfo2_324": (lambda: st.fluorite_hfo2((3, 3, 3), seed=21), 2, 120, 25),
    "hfo2_324_E32": (lambda: st.fluorite_hfo2((3, 3, 3), seed=21), 2, 120, 32),
    "sic_512_E16": (lambda: st.zincblende_sic((4, 4, 4), seed=4), 2, 100, 16),
    "lgps_300_T4": (lambda: st.lgps_frames(nframes=1, nmin=300, nmax=300, seed=3)[0], 4, 150, 25),
}


@pytest.mark.parametrize("name", list(STORE_CASES))
def test_store_mode_matches_oracle_and_recompute(dev, name):
    """Store mode (the forward keeps G and dG/ds per neighbor, the backward reads them) against the
    FP64 oracle at the recomputing pair's tolerances (FP32: 1e-6 |E|, 1e-5 eV/A per component, virial
    1e-5 relative; BF16: force RMSE 1e-3 eV/A, 1e-3 |E|), and against the recomputing pair itself:
    identical energies (the forward arithmetic is the same), FP32 forces within 2e-6 eV/A, BF16 forces
    within RMSE 2e-4 eV/A (G and dG/ds are kept as fp16)."""
    maker, T, m, E = STORE_CASES[name]
    f = maker()
    w = _setup(f, T, m, seed=7, E=E)
    ref = oracle.dp_evaluate(f, w)
    for precision in ("fp32", "bf16"):
        print(f"case fused store {name} {precision}", flush=True)
        Ei, F, V = _fused(f, w, precision, dev, store=True)
        Ei0, F0, V0 = _fused(f, w, precision, dev, store=False)
        np.testing.assert_array_equal(Ei, Ei0)
        if precision == "fp32":
            assert abs(Ei.sum() - ref["E"]) <= 1e-6 * abs(ref["E"])
            assert np.abs(F - ref["F"]).max() <= 1e-5, np.abs(F - ref["F"]).max()
            np.testing.assert_allclose(V.T, ref["V"].reshape(3, 3), atol=1e-5 * np.abs(ref["V"]).max())
            assert np.abs(F - F0).max() <= 2e-6, np.abs(F - F0).max()
        else:
            rmse = np.sqrt(((F - ref["F"]) ** 2).mean())
            d0 = np.sqrt(((F - F0) ** 2).mean())
            print(f"{name}: store bf16 force RMSE {rmse:.2e} vs oracle, {d0:.2e} vs recompute", flush=True)
            assert rmse <= 1e-3
            assert abs(Ei.sum() - ref["E"]) <= 1e-3 * abs(ref["E"])
            assert d0 <= 2e-4


def test_store_mode_absent_types_and_frames(dev):
    """Store mode over packed frames whose atom_type lists differ (a two-type model, a frame with
    one type only): per-frame energies and forces against the FP64 oracle of each frame evaluated
    with that frame's list, FP32 within 1e-6 |E| and 1e-5 eV/A."""
    from matpl_amd.neighbor import find_neighbore_batch
    frames = [st.fluorite_hfo2((2, 2, 2), seed=3), st.fcc_cu((3, 3, 3), seed=1)]
    T, m = 2, 120
    w = _setup(frames[0], T, m, seed=11)
    types = [f.types for f in frames]
    types[1] = np.zeros_like(types[1])
    present = np.array([[True, True], [True, False]])
    nl = find_neighbore_batch([f.lattice for f in frames], [f.frac() for f in frames], types, T, m, w.cfg.rcut,
                              want_dR=False, want_shift=True, device=dev)
    sizes = [f.natoms for f in frames]
    fs = torch.as_tensor(np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64), device=dev)
    af = torch.as_tensor(np.repeat(np.arange(len(frames), dtype=np.int32), sizes), device=dev)
    lat = torch.as_tensor(np.stack([np.asarray(f.lattice).reshape(9) for f in frames]), device=dev)
    pos = torch.cat([nl.frac_wrapped[int(fs[k]):int(fs[k + 1])] @ lat[k].reshape(3, 3)
                     for k in range(len(frames))]).contiguous()
    tm = torch.as_tensor(np.concatenate(types).astype(np.int32), device=dev)
    rev, rst = reverse_map(nl, tm, m, frame_start=fs, atom_frame=af)
    assert int(rst.item()) == 0
    model = FusedDP(w, "fp32", dev, store=True)
    Ei, F, V, status = model(pos, lat, tm, nl.list, nl.shift, nl.counts, m, atom_frame=af, frame_start=fs,
                             nframes=len(frames), rev=rev, type_present=present)
    torch.cuda.synchronize()
    assert int(status.item()) == 0
    Ei, F = Ei.double().cpu().numpy(), F.double().cpu().numpy()
    for k, f in enumerate(frames):
        a, b = int(fs[k]), int(fs[k + 1])
        g = st.Frame(f.lattice, f.pos, types[k]) if k == 1 else f
        ref = oracle.dp_evaluate(g, w, type_present=present[k])
        assert abs(Ei[a:b].sum() - ref["E"]) <= 1e-6 * abs(ref["E"]), k
        assert np.abs(F[a:b] - ref["F"]).max() <= 1e-5, (k, np.abs(F[a:b] - ref["F"]).max())
