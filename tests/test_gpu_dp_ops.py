"""GPU parity of the DP operator contract, neighbor search and env mat vs the oracle."""
import numpy as np
import pytest
import torch

import oracle
from matpl_amd import structures as st
from matpl_amd.dpnet import DPConfig, DPWeights

pytestmark = pytest.mark.gpu


def _rand_op_inputs(B, N, S, ng, seed, dtype):
    rng = np.random.default_rng(seed)
    nl = rng.integers(-1, N + ng, size=(B, N, S)).astype(np.int32)
    nl[:, :, -2:] = -1
    dE = rng.normal(size=(B, N, S * 4))
    Rid = rng.normal(size=(B, N, S * 4, 3))
    Rij = rng.normal(size=(B, N, S, 3))
    Fc = rng.normal(size=(B, N, 3))
    return nl, dE, Rid, Rij, Fc


# FP64 results differ from the oracle only by atomic summation order;
# FP32 inputs are the FP64 inputs rounded, compared at FP32 precision.
TOL = {torch.float64: 1e-11, torch.float32: 2e-5}


def test_neighbor_nonfinite_position_raises(dev):
    """A NaN or infinite coordinate is an error of the search (status -4, the coordinate searched as 0
    so that every index stays in range), for the brute-force (small box) and cell-list searches."""
    from matpl_amd.neighbor import find_neighbore_gpu
    for f, T, m in ((st.fcc_cu(), 1, 100), (st.fluorite_hfo2((6, 6, 6)), 2, 120)):
        for bad in (float("nan"), float("inf")):
            frac = f.frac().copy()
            frac[3, 1] = bad
            print(f"case non-finite position {f.natoms} atoms {bad}", flush=True)
            with pytest.raises(ValueError, match="not finite"):
                find_neighbore_gpu(f.lattice, frac, f.types, T, m, 6.0, device=dev)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("B,N,S,ng", [(1, 7, 5, 0), (2, 33, 24, 3), (1, 0, 4, 0), (3, 17, 1, 0)])
def test_force_virial_ops(dev, dtype, B, N, S, ng):
    nl, dE, Rid, Rij, Fc = _rand_op_inputs(B, N, S, ng, 42 + N, dtype)
    if dtype == torch.float32:
        dE, Rid, Rij, Fc = [x.astype(np.float32).astype(np.float64) for x in (dE, Rid, Rij, Fc)]
    t = lambda x: torch.as_tensor(x, device=dev, dtype=dtype)
    nlt = torch.as_tensor(nl, device=dev).unsqueeze(2)
    F = torch.ops.CalcOps_cuda.calculateForce(nlt, t(dE).view(B, N, 1, S * 4), t(Rid).view(B, N, S * 4, 3), t(Fc),
                                              torch.tensor(ng))[0]
    want = oracle.dp_force_op(nl, dE, Rid, Fc, ng)
    scale = max(1.0, np.abs(want).max(initial=0.0))
    assert F.shape == (B, N + ng, 3)
    np.testing.assert_allclose(F.cpu().double().numpy(), want, atol=TOL[dtype] * scale * 10)
    V, AV = torch.ops.CalcOps_cuda.calculateVirial(nlt, t(dE).view(B, N, 1, S * 4), t(Rij), t(Rid).view(B, N, S * 4, 3),
                                                   torch.tensor(ng))
    wV, wAV = oracle.dp_virial_op(nl, dE, Rij, Rid, ng)
    scale = max(1.0, np.abs(wAV).max(initial=0.0))
    np.testing.assert_allclose(AV.cpu().double().numpy(), wAV, atol=TOL[dtype] * scale * 10)
    np.testing.assert_allclose(V.cpu().double().numpy(), wV, atol=TOL[dtype] * max(1.0, np.abs(wV).max(initial=0.0)) * 100)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_force_virial_backward(dev, dtype):
    B, N, S = 2, 19, 12
    nl, dE, Rid, Rij, Fc = _rand_op_inputs(B, N, S, 0, 7, dtype)
    nl[nl >= N] = -1
    t = lambda x: torch.as_tensor(x, device=dev, dtype=dtype)
    nlt = torch.as_tensor(nl, device=dev).unsqueeze(2)
    dEt = t(dE).view(B, N, 1, S * 4).requires_grad_()
    F = torch.ops.CalcOps_cuda.calculateForce(nlt, dEt, t(Rid).view(B, N, S * 4, 3), t(Fc), torch.tensor(0))[0]
    gF = np.random.default_rng(1).normal(size=(B, N, 3))
    (g,) = torch.autograd.grad(F, dEt, t(gF))
    want = oracle.dp_force_grad_op(nl, Rid, gF, 0)
    np.testing.assert_allclose(g.detach().cpu().double().numpy().reshape(want.shape), want,
                               atol=TOL[dtype] * 100)
    V = torch.ops.CalcOps_cuda.calculateVirial(nlt, dEt, t(Rij), t(Rid).view(B, N, S * 4, 3), torch.tensor(0))[0]
    gV = np.random.default_rng(2).normal(size=(B, 9))
    (g,) = torch.autograd.grad(V, dEt, t(gV))
    want = oracle.dp_virial_grad_op(nl, Rij, Rid, gV)
    np.testing.assert_allclose(g.detach().cpu().double().numpy().reshape(want.shape), want,
                               atol=TOL[dtype] * 100)


FRAMES = {
    "cu108": (lambda: st.fcc_cu(), 1, 100, 6.0),                      # small box: L < 2 rc
    "cu_2x2x2": (lambda: st.fcc_cu((2, 2, 2), seed=3), 1, 300, 6.0),  # many images per j
    "sic_4x4x4": (lambda: st.zincblende_sic((4, 4, 4)), 2, 100, 6.0),  # 17.4 A: cell list, nc = 2
    "hfo2_5x5x5": (lambda: st.fluorite_hfo2((5, 5, 5)), 2, 120, 6.0),  # cell list, nc = 4
    "hfo2_skin": (lambda: st.fluorite_hfo2((5, 5, 6)), 2, 150, 7.0),
}


@pytest.mark.parametrize("name", list(FRAMES))
def test_neighbor_bit_exact(dev, name):
    from matpl_amd.neighbor import find_neighbore_gpu
    maker, T, m, rc = FRAMES[name]
    f = maker()
    frac = f.frac()
    lst, dR = oracle.dp_find_neighbor(f.lattice, frac, f.types, T, m, rc)
    nl = find_neighbore_gpu(f.lattice, frac, f.types, T, m, rc, device=dev, want_shift=True)
    assert torch.equal(nl.list.cpu(), torch.as_tensor(lst))
    assert np.array_equal(nl.dR.cpu().numpy(), dR)                       # bit-exact FP64
    np.testing.assert_array_equal(nl.counts.cpu().numpy(),
                                  np.stack([(lst[:, t * m:(t + 1) * m] > 0).sum(1) for t in range(T)], 1))
    # the packed image shift reproduces dR from the wrapped fractions
    fw = nl.frac_wrapped.cpu().numpy()
    sh = nl.shift.cpu().numpy()
    s = np.stack([(sh & 255) - 128, ((sh >> 8) & 255) - 128, ((sh >> 16) & 255) - 128], -1)
    ii, kk = np.nonzero(lst)
    jj = lst[ii, kk] - 1
    rec = (fw[jj] - fw[ii] + s[ii, kk]) @ f.lattice
    np.testing.assert_allclose(rec, dR[ii, kk], atol=1e-9)


def test_neighbor_cell_vs_brute(dev):
    from matpl_amd.neighbor import find_neighbore_gpu
    f = st.fluorite_hfo2((6, 6, 6))
    a = find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, 120, 6.0, device=dev)
    b = find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, 120, 6.0, device=dev, force_brute=True)
    assert torch.equal(a.list, b.list) and torch.equal(a.dR, b.dR)


def test_neighbor_batch_matches_oracle(dev):
    from matpl_amd.neighbor import find_neighbore_batch
    frames = st.lgps_frames(nframes=6, nmin=40, nmax=120, seed=9)
    T, m, rc = 4, 150, 6.0
    nl = find_neighbore_batch([f.lattice for f in frames], [f.frac() for f in frames],
                              [f.types for f in frames], T, m, rc, device=dev)
    off = 0
    for f in frames:
        lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, T, m, rc)
        assert torch.equal(nl.list[off:off + f.natoms].cpu(), torch.as_tensor(lst))
        assert np.array_equal(nl.dR[off:off + f.natoms].cpu().numpy(), dR)
        off += f.natoms


def test_neighbor_overflow_raises(dev):
    from matpl_amd.neighbor import find_neighbore_gpu
    f = st.fluorite_hfo2((5, 5, 5))
    with pytest.raises(OverflowError):
        find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, 20, 6.0, device=dev)
    f = st.fcc_cu()
    with pytest.raises(OverflowError):
        find_neighbore_gpu(f.lattice, f.frac(), f.types, 1, 50, 6.0, device=dev)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_env_mat_and_calc_stat(dev, dtype):
    from matpl_amd.neighbor import calc_stat, env_mat, find_neighbore_gpu, image_dr
    f = st.fluorite_hfo2((4, 4, 4))
    T, m = 2, 120
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, T, m, 6.0)
    davg, dstd = oracle.dp_calc_stat(lst, dR, f.types, T, 5.5, 6.0)
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, T, m, 6.0, device=dev)
    tm = torch.as_tensor(f.types, device=dev)
    gdavg, gdstd = calc_stat(nl.list, nl.dR, tm, T, 5.5, 6.0)
    np.testing.assert_allclose(gdavg.cpu().numpy(), davg, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(gdstd.cpu().numpy(), dstd, rtol=1e-12)
    idr = oracle.dp_image_dr(lst, dR)
    Ri, Ri_d = oracle.dp_env_mat(idr, f.types, davg, dstd, 5.5, 6.0)
    S = T * m
    full = lambda a: torch.as_tensor(np.tile(a[:, None, :], (1, S, 1)).reshape(T, -1), device=dev)
    gRi, gRid = env_mat(image_dr(nl.list, nl.dR).to(dtype), tm, full(davg), full(dstd), 5.5, 6.0)
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    np.testing.assert_allclose(gRi[0].cpu().double().numpy(), Ri, atol=tol * np.abs(Ri).max())
    np.testing.assert_allclose(gRid[0].cpu().double().numpy(), Ri_d, atol=tol * np.abs(Ri_d).max())


def test_dp_torch_fp64_matches_oracle(dev):
    """The plain FP64 drop-in path (torch nets + HIP ops) equals the oracle."""
    from matpl_amd.dp import DPTorch
    from matpl_amd.neighbor import find_neighbore_gpu, image_dr
    f = st.fluorite_hfo2((3, 3, 3), seed=21)
    cfg = DPConfig(ntypes=2, m=120)
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 2, 120, 6.0)
    davg, dstd = oracle.dp_calc_stat(lst, dR, f.types, 2, 5.5, 6.0)
    w = DPWeights.random(cfg, seed=5, davg=davg, dstd=dstd)
    ref = oracle.dp_evaluate(f, w)
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, 120, 6.0, device=dev)
    model = DPTorch(w, torch.float64, dev)
    Etot, Ei, F, V = model(nl.list.unsqueeze(0), torch.as_tensor(f.types, device=dev),
                           image_dr(nl.list, nl.dR).unsqueeze(0))
    assert abs(Etot.item() - ref["E"]) < 1e-9 * abs(ref["E"])
    np.testing.assert_allclose(Ei[0].detach().cpu().numpy(), ref["Ei"], atol=1e-10)
    np.testing.assert_allclose(F[0].detach().cpu().numpy(), ref["F"], atol=1e-10)
    # GPU op convention V[a][c] = sum f_a dR_c is the transpose of the CPU branch's
    np.testing.assert_allclose(V[0].detach().cpu().numpy().reshape(3, 3).T, ref["V"].reshape(3, 3),
                               atol=1e-8)
