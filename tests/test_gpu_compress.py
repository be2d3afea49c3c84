"""calculateCompress on the GPU against the oracle (oracle/compress.py), and the compressed DP model
against the uncompressed one on the same structure."""
import numpy as np
import pytest
import torch

import oracle
import oracle.compress as oc
from matpl_amd import structures as st
from matpl_amd.dpnet import DPConfig, DPWeights

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K", [4, 6])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_compress_op_matches_oracle(dev, K, dtype):
    rng = np.random.default_rng(K)
    n, L = 1000, 25
    f2 = rng.uniform(-0.01, 0.1, size=(n, 1))
    coef = rng.normal(size=(n, L, K))
    gG = rng.normal(size=(n, L))
    want, want_g = oc.compress_forward(f2, coef, gG)
    f2_t = torch.as_tensor(f2, dtype=dtype, device=dev).requires_grad_()
    G = torch.ops.CalcOps_cuda.calculateCompress(f2_t, torch.as_tensor(coef, dtype=dtype, device=dev))[0]
    (g,) = torch.autograd.grad(G, f2_t, torch.as_tensor(gG, dtype=dtype, device=dev))
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    np.testing.assert_allclose(G.detach().cpu().double().numpy(), want, rtol=tol, atol=tol * np.abs(want).max())
    np.testing.assert_allclose(g.cpu().double().numpy()[:, 0], want_g, rtol=tol,
                               atol=tol * np.abs(want_g).max())
    # empty input
    e = torch.ops.CalcOps_cuda.calculateCompress(torch.zeros((0, 1), dtype=dtype, device=dev),
                                                 torch.zeros((0, L, K), dtype=dtype, device=dev))[0]
    assert e.shape == (0, L)


@pytest.mark.parametrize("order", [3, 5])
def test_compressed_model_matches_uncompressed(dev, order):
    from matpl_amd.compress import build_table
    from matpl_amd.dp import DPTorch
    from matpl_amd.neighbor import find_neighbore_gpu, image_dr
    f = st.fluorite_hfo2((3, 3, 3), seed=23)
    cfg = DPConfig(ntypes=2, m=120)
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 2, 120, 6.0)
    davg, dstd = oracle.dp_calc_stat(lst, dR, f.types, 2, 5.5, 6.0)
    w = DPWeights.random(cfg, seed=9, davg=davg, dstd=dstd)
    ref = oracle.dp_evaluate(f, w)
    plain = DPTorch(w, torch.float64, dev)
    s_max_raw = 1.0 / 1.5          # s(r) = 1/r below Rmin; shortest bond here is above 1.5 A
    ct = build_table(list(plain.emb), davg, dstd, s_max_raw, dx=0.001, order=order, device=dev)
    comp = DPTorch(w, torch.float64, dev, compress=ct)
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, 120, 6.0, device=dev)
    args = (nl.list.unsqueeze(0), torch.as_tensor(f.types, device=dev), image_dr(nl.list, nl.dR).unsqueeze(0))
    E, Ei, F, V = comp(*args, create_graph=False)
    assert abs(E.item() - ref["E"]) < 1e-7 * abs(ref["E"])
    f_err = np.abs(F[0].detach().cpu().numpy() - ref["F"]).max()
    assert f_err < 1e-6, f_err


@pytest.mark.parametrize("order", [3, 5])
def test_fused_compressed_matches_compressed_model(dev, order):
    """Fused kernels in table mode (G and dG/ds from the table by Horner) against DPTorch on the same
    table in FP64: FP32 fused within 1e-6 |E|, 1e-5 eV/A per force component, 1e-5 relative virial;
    BF16 fused within force RMSE 1e-3 eV/A of the uncompressed FP64 oracle."""
    from matpl_amd.compress import build_table
    from matpl_amd.dp import DPTorch
    from matpl_amd.dp_fused import FusedDP
    from matpl_amd.neighbor import find_neighbore_gpu, image_dr, reverse_map
    f = st.fluorite_hfo2((3, 3, 3), seed=23)
    cfg = DPConfig(ntypes=2, m=120)
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 2, 120, 6.0)
    davg, dstd = oracle.dp_calc_stat(lst, dR, f.types, 2, 5.5, 6.0)
    w = DPWeights.random(cfg, seed=9, davg=davg, dstd=dstd)
    ref = oracle.dp_evaluate(f, w)
    plain = DPTorch(w, torch.float64, dev)
    ct = build_table(list(plain.emb), davg, dstd, 1.0 / 1.5, dx=0.01, order=order, device=dev)
    comp = DPTorch(w, torch.float64, dev, compress=ct)
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, 120, 6.0, device=dev, want_shift=True)
    tm = torch.as_tensor(f.types, device=dev)
    E0, _, F0, V0 = comp(nl.list.unsqueeze(0), tm, image_dr(nl.list, nl.dR).unsqueeze(0), create_graph=False)
    E0, F0, V0 = E0.item(), F0[0].detach().cpu().numpy(), V0[0].detach().cpu().numpy()
    L = torch.as_tensor(f.lattice, device=dev)
    pos = (nl.frac_wrapped @ L).contiguous()
    rev, rst = reverse_map(nl, tm, 120)
    assert int(rst.item()) == 0
    for precision in ("fp32", "bf16"):
        print(f"case fused compressed order {order} {precision}", flush=True)
        model = FusedDP(w, precision, dev, compress=ct)
        Ei, F, V, status = model(pos, L.reshape(1, 9).contiguous(), tm, nl.list, nl.shift, nl.counts, 120, rev=rev)
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        E = float(Ei.double().sum().item())
        F = F.double().cpu().numpy()
        V = V[0].cpu().numpy()
        if precision == "fp32":
            assert abs(E - E0) < 1e-6 * abs(E0), (E, E0)
            assert np.abs(F - F0).max() < 1e-5
            assert np.abs(V - V0).max() < 1e-5 * np.abs(V0).max()
        else:
            rmse = float(np.sqrt(((F - ref["F"]) ** 2).mean()))
            assert rmse < 1e-3, rmse
            assert abs(E - ref["E"]) < 1e-3 * abs(ref["E"])
