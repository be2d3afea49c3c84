"""NEP operators on the GPU against the FP64 oracle (oracle/nep_oracle.c, oracle/nep.py).

Neighbor lists and counts bit-exact; features, forces and virials to 1e-10
relative in float64 and 2e-5 in float32; first- and second-order operator
gradients by gradcheck / gradgradcheck; the force-loss parameter gradient of
the whole model against the oracle's torch double backward.
"""
import numpy as np
import pytest
import torch

import oracle.nep as on
from matpl_amd import nep as mn
from matpl_amd import nep_ops
from matpl_amd import structures as st
from matpl_amd.nepnet import NEPConfig, NEPWeights

pytestmark = pytest.mark.gpu


def _frames():
    return [st.zincblende_sic((2, 2, 2), seed=4), st.agausic_frames(nframes=2, nmin=30, nmax=50, seed=5)[0],
            st.zincblende_sic((3, 3, 3), seed=3)]


def _rel(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max(initial=0.0) / max(np.abs(b).max(initial=0.0), 1e-30))


def test_neighbor_bit_exact(dev):
    frames = _frames()
    pk = on.pack_frames(frames, 6.0)
    want = on.neighbor(pk, 6.0, 4.0, 200, 60)
    gp = mn.pack_frames(frames, 6.0, dev)
    for k in ("box", "box_orig", "num_cell"):
        assert np.array_equal(gp[k].cpu().numpy().reshape(-1), np.asarray(pk[k]).reshape(-1)), k
    got = nep_ops.calculate_neighbor(gp["num_atoms"], gp["atom_type_map"], gp["atom_type_map"], gp["box"],
                                     gp["box_orig"], gp["num_cell"], gp["position"], 6.0, 4.0, 200, 60, True)
    for name, t in zip(("NN_r", "NN_a", "NL_r", "NL_a", "Ri_r", "Ri_a"), got):
        assert np.array_equal(t.cpu().numpy(), want[name]), name
    # without rij: [x, y, z]
    got3 = nep_ops.calculate_neighbor(gp["num_atoms"], gp["atom_type_map"], gp["atom_type_map"], gp["box"],
                                      gp["box_orig"], gp["num_cell"], gp["position"], 6.0, 4.0, 200, 60, False)
    assert np.array_equal(got3[4].cpu().numpy(), want["Ri_r"][..., 1:])
    for with_type in (False, True):
        r, a = on.maxneigh(pk, 6.0, 4.0, 4, with_type)
        g = nep_ops.calculate_maxneigh(gp["num_atoms"], gp["box"], gp["box_orig"], gp["num_cell"], gp["position"],
                                       6.0, 4.0, 4, gp["atom_type_map"], with_type)
        assert np.array_equal(g[0].cpu().numpy(), r) and np.array_equal(g[1].cpu().numpy(), a)


def test_neighbor_bit_exact_larger_frames(dev):
    """Frames of 512 and 640 atoms in boxes thicker than 2 rc (one replica: one pass counts and writes,
    8 and 10 chunks of 64 candidates per atom); lists and Ri bit-identical to the oracle."""
    frames = [st.zincblende_sic((4, 4, 4), seed=6), st.zincblende_sic((5, 4, 4), seed=7)]
    pk = on.pack_frames(frames, 6.0)
    want = on.neighbor(pk, 6.0, 4.0, 200, 60)
    gp = mn.pack_frames(frames, 6.0, dev)
    assert int(gp["num_cell"].max()) == 1
    got = nep_ops.calculate_neighbor(gp["num_atoms"], gp["atom_type_map"], gp["atom_type_map"], gp["box"],
                                     gp["box_orig"], gp["num_cell"], gp["position"], 6.0, 4.0, 200, 60, True)
    for name, t in zip(("NN_r", "NN_a", "NL_r", "NL_a", "Ri_r", "Ri_a"), got):
        assert np.array_equal(t.cpu().numpy(), want[name]), name


def test_neighbor_float_rows_equal_cast(dev):
    """calculate_neighbor(ri_dtype=float32): the same lists and counts, and Ri rows equal to the
    float64 rows cast to float (each value rounded once), for one-replica and replicated frames."""
    frames = _frames() + [st.zincblende_sic((4, 4, 4), seed=6)]
    gp = mn.pack_frames(frames, 6.0, dev)
    args = (gp["num_atoms"], gp["atom_type_map"], gp["atom_type_map"], gp["box"], gp["box_orig"], gp["num_cell"],
            gp["position"], 6.0, 4.0, 200, 60, True)
    d = nep_ops.calculate_neighbor(*args)
    f = nep_ops.calculate_neighbor(*args, ri_dtype=torch.float32)
    for k in range(4):
        assert torch.equal(d[k], f[k])
    for k in (4, 5):
        assert f[k].dtype == torch.float32 and torch.equal(f[k], d[k].float())


def test_neighbor_overflow_raises(dev):
    gp = mn.pack_frames([st.zincblende_sic((2, 2, 2), seed=4)], 6.0, dev)
    with pytest.raises(OverflowError):
        nep_ops.calculate_neighbor(gp["num_atoms"], gp["atom_type_map"], gp["atom_type_map"], gp["box"],
                                   gp["box_orig"], gp["num_cell"], gp["position"], 6.0, 4.0, 20, 60, True)


@pytest.mark.parametrize("lmax", [(4, 2, 0), (4, 2, 1), (4, 0, 0), (2, 0, 0)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_features_match_oracle(dev, lmax, dtype):
    frames = _frames()
    cfg = NEPConfig(ntypes=4, l_max=list(lmax))
    w = NEPWeights.random(cfg, seed=11)
    pk = on.pack_frames(frames, 6.0)
    nb = on.neighbor(pk, 6.0, 4.0, 200, 60)
    am = pk["atom_type_map"]
    qr = on.radial_feats(w.c2, torch.as_tensor(nb["Ri_r"]), nb["NL_r"], am, 6.0).numpy()
    qa = on.angular_feats(w.c3, torch.as_tensor(nb["Ri_a"]), nb["NL_a"], am, 4.0, *lmax).numpy()
    T = lambda a, dt=dtype: torch.as_tensor(a, dtype=dt, device=dev)
    gr = nep_ops.calculateNepFeat(T(w.c2), T(nb["Ri_r"]), T(nb["NL_r"], torch.int64), T(am, torch.int64), None,
                                  6.0, 0)[0]
    ga = nep_ops.calculateNepMbFeat(T(w.c3), T(nb["Ri_a"]), T(nb["NL_a"], torch.int64), T(am, torch.int64), None,
                                    cfg.nr, *lmax, 4.0)[0]
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    assert gr.dtype == dtype and ga.dtype == dtype
    assert _rel(gr.cpu(), qr) < tol
    assert _rel(ga.cpu(), qa) < tol


def test_descriptor_matches_oracle(dev):
    frames = _frames()[:2]
    cfg = NEPConfig(ntypes=4, l_max=[4, 2, 1])
    w = NEPWeights.random(cfg, seed=12)
    pk = on.pack_frames(frames, 6.0)
    r = on.nep_evaluate(frames, w)
    gp = mn.pack_frames(frames, 6.0, dev)
    nb = nep_ops.calculate_neighbor(gp["num_atoms"], gp["atom_type_map"], gp["atom_type_map"], gp["box"],
                                    gp["box_orig"], gp["num_cell"], gp["position"], 6.0, 6.0, 200, 200, False)
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev)
    q = nep_ops.calculate_descriptor(T(w.c2), T(w.c3), nb[4], nb[2], gp["atom_type_map"], 6.0, 4.0, 200, 4, 2, 1)[0]
    assert _rel(q.cpu(), r["q"]) < 1e-11


def _rand_inputs(dev, N=6, S=7, T=2, nm=3, nb=5, angular=False, seed=0):
    g = torch.Generator().manual_seed(seed)
    NL = torch.randint(0, N, (N, S), generator=g)
    cnt = torch.randint(S // 2, S + 1, (N,), generator=g)
    NL[torch.arange(S)[None, :] >= cnt[:, None]] = -1
    xyz = torch.randn(N, S, 3, generator=g, dtype=torch.float64)
    xyz = xyz / xyz.norm(dim=-1, keepdim=True) * (1.0 + 2.5 * torch.rand(N, S, 1, generator=g, dtype=torch.float64))
    r = xyz.norm(dim=-1, keepdim=True) * (1 + 0.01 * torch.randn(N, S, 1, generator=g, dtype=torch.float64))
    Ri = torch.cat([r, xyz], -1)
    Ri[NL < 0] = 0
    am = torch.randint(0, T, (N,), generator=g)
    c = torch.randn(T, T, nm, nb, generator=g, dtype=torch.float64)
    return (c.to(dev), Ri.to(dev), NL.to(dev), am.to(dev))


def test_radial_gradcheck(dev):
    c, Ri, NL, am = _rand_inputs(dev)
    c.requires_grad_(); Ri.requires_grad_()
    f = lambda c_, R_: nep_ops.NepFeatRadial.apply(c_, R_, NL, am, 4.0)
    assert torch.autograd.gradcheck(f, (c, Ri), eps=1e-6, atol=1e-7, rtol=1e-6)
    # second order as force training uses it: d(gRi)/d(upstream feature grad, c)
    g = torch.randn(NL.shape[0], c.shape[2], dtype=torch.float64, device=dev, requires_grad=True)
    h = lambda g_, c_: nep_ops.NepFeatRadialGrad.apply(g_, c_, Ri.detach(), NL, am, 4.0)[1]
    assert torch.autograd.gradcheck(h, (g, c), eps=1e-6, atol=1e-7, rtol=1e-6)


@pytest.mark.parametrize("lmax", [(4, 2, 1), (4, 2, 0), (3, 0, 0)])
def test_angular_gradcheck(dev, lmax):
    c, Ri, NL, am = _rand_inputs(dev, seed=1)
    c.requires_grad_(); Ri.requires_grad_()
    f = lambda c_, R_: nep_ops.NepFeatAngular.apply(c_, R_, NL, am, 4.0, *lmax)
    assert torch.autograd.gradcheck(f, (c, Ri), eps=1e-6, atol=1e-7, rtol=1e-6)
    nf = nep_ops._nangular(c.shape[2], *lmax)
    g = torch.randn(NL.shape[0], nf, dtype=torch.float64, device=dev, requires_grad=True)

    def h(g_, c_):
        R = Ri.detach()
        s = torch.empty((NL.shape[0], c.shape[2] * 24), dtype=torch.float64, device=dev)
        nep_ops._feat_call(0, True, c_.detach(), R, NL, am, 4.0, lmax, feats=torch.empty(NL.shape[0], nf,
                           dtype=torch.float64, device=dev), s_save=s)
        return nep_ops.NepFeatAngularGrad.apply(g_, c_, R, NL, am, s, 4.0, *lmax)[1]

    # the saved s sums are recomputed for every perturbed c, so the check sees d/dc through them too
    assert torch.autograd.gradcheck(h, (g, c), eps=1e-6, atol=1e-7, rtol=1e-6)


def test_force_virial_ops_and_grads(dev):
    _, Ri, NL, _ = _rand_inputs(dev, N=9, S=8, seed=2)
    N, S = NL.shape
    g = torch.Generator().manual_seed(3)
    dE = torch.randn(N, S, 4, generator=g, dtype=torch.float64).to(dev) * (NL >= 0).unsqueeze(-1)
    Rid = mn.ri_d(Ri)
    num_atom = torch.tensor([4, 5], device=dev)
    centre = -torch.matmul(dE.reshape(N, 1, -1), Rid.reshape(N, -1, 3)).squeeze(-2)
    F = nep_ops.calculateNepForce(NL, dE, Rid, centre)[0]
    V, AV = nep_ops.calculateNepVirial(NL, dE, Ri[..., 1:].contiguous(), Rid, num_atom)
    # direct restatement on the CPU
    dEc, Ridc, NLc, Ric = dE.cpu(), Rid.cpu(), NL.cpu(), Ri.cpu()
    pair = torch.einsum("nsk,nska->nsa", dEc, Ridc)
    Fw = -pair.sum(1)
    AVw = torch.zeros(N, 9, dtype=torch.float64)
    for i in range(N):
        for s in range(S):
            j = int(NLc[i, s])
            if j < 0:
                continue
            Fw[j] += pair[i, s]
            AVw[j] += torch.outer(pair[i, s], Ric[i, s, 1:]).reshape(9)
    Vw = torch.stack([AVw[:4].sum(0), AVw[4:].sum(0)])
    assert torch.allclose(F.cpu(), Fw, rtol=1e-12, atol=1e-12)
    assert torch.allclose(AV.cpu(), AVw, rtol=1e-12, atol=1e-12)
    assert torch.allclose(V.cpu(), Vw, rtol=1e-12, atol=1e-12)
    # backward wrt dE: the op's gradient covers centre + neighbor terms
    dE.requires_grad_()
    f = lambda d: nep_ops.NepForce.apply(NL, d, Rid, torch.zeros(N, 3, dtype=torch.float64, device=dev)) - \
        torch.matmul(d.reshape(N, 1, -1), Rid.reshape(N, -1, 3)).squeeze(-2).detach()
    gF = torch.randn(N, 3, dtype=torch.float64, device=dev)
    (got,) = torch.autograd.grad(f(dE), dE, gF)
    want = torch.zeros(N, S, 4, dtype=torch.float64)
    for i in range(N):
        for s in range(S):
            j = int(NLc[i, s])
            if j < 0:
                continue
            want[i, s] = Ridc[i, s] @ (gF.cpu()[j] - gF.cpu()[i])
    assert torch.allclose(got.cpu(), want, rtol=1e-12, atol=1e-12)
    gV = torch.randn(2, 9, dtype=torch.float64, device=dev)
    (gotv,) = torch.autograd.grad(nep_ops.NepVirial.apply(NL, dE, Ri[..., 1:].contiguous(), Rid, num_atom)[0], dE,
                                  gV)
    wantv = torch.zeros(N, S, 4, dtype=torch.float64)
    for i in range(N):
        fr = 0 if i < 4 else 1
        for s in range(S):
            if int(NLc[i, s]) < 0:
                continue
            wantv[i, s] = Ridc[i, s] @ gV.cpu()[fr].reshape(3, 3) @ Ric[i, s, 1:]
    assert torch.allclose(gotv.cpu(), wantv, rtol=1e-12, atol=1e-12)
    # model-path variants: centre term by HIP reduction, frame virial without per-atom rows
    assert torch.allclose(nep_ops.nep_centre_force(dE, Rid).cpu(), -pair.sum(1), rtol=1e-12, atol=1e-12)
    # (sums by the centre atom's frame; the random lists here cross frames, real ones never do)
    Vc = torch.zeros(2, 9, dtype=torch.float64)
    for i in range(N):
        for s in range(S):
            if int(NLc[i, s]) >= 0:
                Vc[0 if i < 4 else 1] += torch.outer(pair[i, s], Ric[i, s, 1:]).reshape(9)
    Vf = nep_ops.nep_frame_virial(NL, dE, Ri[..., 1:].contiguous(), Rid, num_atom)
    assert torch.allclose(Vf.cpu(), Vc, rtol=1e-12, atol=1e-12)
    (gotf,) = torch.autograd.grad(Vf, dE, gV)
    assert torch.allclose(gotf.cpu(), wantv, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("N", [600, 2000])
def test_force_virial_ops_wide_lists(dev, N):
    """calculateNepForce / calculateNepVirial on lists naming atoms anywhere in the batch: per block of
    atoms the named range is up to N wide, so the LDS copies hold it (600 atoms) or, for the virial at
    2,000 atoms, exceed their 64 KB and the block adds to global memory. Against index_add on the CPU;
    FP64 within 1e-12, FP32 within 1e-5 relative."""
    S = 16
    g = torch.Generator().manual_seed(N)
    NL = torch.randint(0, N, (N, S), generator=g)
    NL[torch.rand(N, S, generator=g) < 0.2] = -1
    Ri = torch.cat([torch.rand(N, S, 1, generator=g, dtype=torch.float64) + 1.0,
                    torch.randn(N, S, 3, generator=g, dtype=torch.float64)], -1)
    Rid = mn.ri_d(Ri)
    dE = torch.randn(N, S, 4, generator=g, dtype=torch.float64)
    F0 = torch.randn(N, 3, generator=g, dtype=torch.float64)
    m = NL >= 0
    pair = torch.einsum("nsk,nska->nsa", dE, Rid)[m]
    j = NL[m]
    Fw = F0.clone().index_add_(0, j, pair)
    AVw = torch.zeros(N, 9, dtype=torch.float64).index_add_(
        0, j, (pair.unsqueeze(-1) * Ri[..., 1:][m].unsqueeze(-2)).reshape(-1, 9))
    print(f"case wide lists N={N}", flush=True)
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        T = lambda x: x.to(dt).to(dev)
        NLd = NL.to(dev)
        F = nep_ops.calculateNepForce(NLd, T(dE), T(Rid), T(F0))[0]
        V, AV = nep_ops.calculateNepVirial(NLd, T(dE), T(Ri[..., 1:].contiguous()), T(Rid),
                                           torch.tensor([N], device=dev))
        for got, want in ((F, Fw), (AV, AVw), (V, AVw.sum(0, keepdim=True))):
            err = (got.double().cpu() - want).abs().max().item()
            assert err <= tol * want.abs().max().item(), (dt, err)


def test_frame_virial_many_frames(dev):
    """Frames shorter and longer than one wavefront's run of atoms, and an empty frame."""
    _, Ri, NL, _ = _rand_inputs(dev, N=40, S=70, seed=7)
    N, S = NL.shape
    g = torch.Generator().manual_seed(8)
    sizes = [3, 0, 1, 17, 9, 10]
    num_atom = torch.tensor(sizes, device=dev)
    # neighbor indices inside each atom's own frame, -1 padding kept
    start = np.cumsum([0] + sizes)
    fr = np.repeat(np.arange(len(sizes)), sizes)
    lo = torch.as_tensor(start[fr]).unsqueeze(1)
    width = torch.as_tensor(np.array(sizes)[fr]).unsqueeze(1)
    NL = torch.where(NL.cpu() >= 0, lo + torch.randint(0, 1 << 20, (N, S), generator=g) % width, NL.cpu()).to(dev)
    dE = torch.randn(N, S, 4, generator=g, dtype=torch.float64).to(dev)
    Rid = mn.ri_d(Ri)
    AV = nep_ops.calculateNepVirial(NL, dE, Ri[..., 1:].contiguous(), Rid, num_atom)[1].cpu()
    want = torch.stack([r.sum(0) for r in AV.split(num_atom.tolist())])
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 1e-4)):
        got = nep_ops.nep_frame_virial(NL, dE.to(dt), Ri[..., 1:].to(dt).contiguous(), Rid.to(dt), num_atom)
        assert torch.allclose(got.double().cpu(), want, rtol=tol, atol=tol * want.abs().max().item())


def _framed_lists(dev, sizes, S, seed):
    """Random lists whose entries stay in their atom's frame (-1 padded at the tail), with Ri rows
    [r, x, y, z] and one used entry per frame of 2+ atoms at |r| below 1e-5 (its Ri_d is zero)."""
    g = torch.Generator().manual_seed(seed)
    N = sum(sizes)
    start = np.cumsum([0] + sizes)
    fr = np.repeat(np.arange(len(sizes)), sizes)
    lo = torch.as_tensor(start[fr]).unsqueeze(1)
    width = torch.as_tensor(np.maximum(np.array(sizes)[fr], 1)).unsqueeze(1)
    NL = lo + torch.randint(0, 1 << 20, (N, S), generator=g) % width
    cnt = torch.randint(0, S + 1, (N,), generator=g)
    NL[torch.arange(S)[None, :] >= cnt[:, None]] = -1
    xyz = torch.randn(N, S, 3, generator=g, dtype=torch.float64) * 2.0
    Ri = torch.cat([xyz.norm(dim=-1, keepdim=True), xyz], -1)
    for f in range(len(sizes)):
        if sizes[f] >= 2 and cnt[start[f]] > 0:
            Ri[start[f], 0] = torch.tensor([5e-6, 3e-6, 4e-6, 0.0], dtype=torch.float64)
    Ri[NL < 0] = 0
    dE = torch.randn(N, S, 4, generator=g, dtype=torch.float64) * (NL >= 0).unsqueeze(-1)
    return NL.to(dev), Ri.to(dev), dE.to(dev), torch.tensor(sizes, device=dev)


@pytest.mark.parametrize("sizes,S", [([3, 0, 1, 70, 200, 9, 64, 65], 24), ([3000, 5, 40], 12),
                                     ([420, 3, 0, 77] * 10 + [420] * 30, 8)])
def test_fused_force_virial_matches_op_sequence(dev, sizes, S):
    """nep_force_virial (Ri_d formed in the kernel) against the reference's sequence ri_d -> centre
    term -> calculateNepForce -> calculateNepVirial frame sums, values and dE gradients: few atoms or
    a large frame (global atomics per entry), and 17,600 atoms in frames of up to 420 (per-block LDS
    copies of F and their reduction). FP64 within 1e-12 relative, FP32 within 1e-5."""
    NL, Ri, dE, num_atom = _framed_lists(dev, sizes, S, seed=len(sizes))
    N = NL.shape[0]
    print(f"case fused force/virial sizes={sizes} S={S}", flush=True)
    gF = torch.randn(N, 3, dtype=torch.float64, device=dev)
    gV = torch.randn(len(sizes), 9, dtype=torch.float64, device=dev)

    def seq(d_):
        # the op's backward carries the centre term, so the centre force enters as a constant
        Rid = mn.ri_d(Ri)
        centre = -torch.matmul(d_.detach().reshape(N, 1, -1), Rid.reshape(N, -1, 3)).squeeze(-2)
        F = nep_ops.NepForce.apply(NL, d_, Rid, centre)
        V = nep_ops.nep_frame_virial(NL, d_, Ri[..., 1:].contiguous(), Rid, num_atom)
        return F, V

    d64 = dE.clone().requires_grad_()
    Fw, Vw = seq(d64)
    gw = torch.autograd.grad((Fw * gF).sum() + (Vw * gV).sum(), d64)[0]
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        d = dE.to(dt).clone().requires_grad_()
        F, V = nep_ops.nep_force_virial(NL, d, Ri.to(dt), num_atom)
        g = torch.autograd.grad((F * gF.to(dt)).sum() + (V * gV.to(dt)).sum(), d)[0]
        for got, want in ((F, Fw), (V, Vw), (g, gw)):
            err = (got.double() - want).abs().max().item()
            assert err <= tol * max(want.abs().max().item(), 1.0), (dt, err)
    # upstream gradient of F only, and of V only
    d = dE.clone().requires_grad_()
    F, V = nep_ops.nep_force_virial(NL, d, Ri, num_atom)
    gf_only = torch.autograd.grad((F * gF).sum(), d, retain_graph=True)[0]
    gv_only = torch.autograd.grad((V * gV).sum(), d)[0]
    assert torch.allclose(gf_only + gv_only, gw, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("lmax", [(4, 2, 0), (4, 2, 1)])
def test_model_energy_force_virial(dev, lmax):
    frames = _frames()
    cfg = NEPConfig(ntypes=4, l_max=list(lmax))
    w = NEPWeights.random(cfg, seed=13)
    want = on.nep_evaluate(frames, w)
    gp = mn.pack_frames(frames, cfg.rcut_radial, dev)
    nb = mn.neighbors(gp, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes)
    model = mn.NEPTorch(w, torch.float64, dev)
    Etot, Ei, F, V = model(nb["NN_r"], nb["NL_r"], nb["Ri_r"], nb["NN_a"], nb["NL_a"], nb["Ri_a"], gp["num_atoms"],
                           gp["atom_type_map"], create_graph=False)
    assert _rel(Etot.detach().cpu().reshape(-1), want["E"]) < 1e-12
    assert _rel(Ei.detach().cpu(), want["Ei"]) < 1e-12
    assert _rel(F.detach().cpu(), want["F"]) < 1e-10
    assert _rel(V.detach().cpu(), want["V"]) < 1e-10


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_trained_hfo2_model(dev, dtype, tol):
    """The trained HfO2 NEP model the reference ships (tests/golden/hfo2_nep; nep4, O/Hf, l_max [4,2,1],
    100 hidden units, rc 6/6 A) on four of its DFT-labelled training frames: the HIP path equals the
    FP64 oracle, and reproduces the DFT forces."""
    import os
    from matpl_amd.nep_io import read_nep_txt
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hfo2_nep")
    w, el, mx = read_nep_txt(os.path.join(here, "nep.txt"))
    d = np.load(os.path.join(here, "frames.npz"))
    types = np.array([{8: el.index("O"), 72: el.index("Hf")}[int(z)] for z in d["image_type"]], np.int32)
    frames = []
    for k in range(d["lattice"].shape[0]):
        L = d["lattice"][k].reshape(3, 3)
        frames.append(st.Frame(L, d["position"][k].reshape(-1, 3) @ L, types, el))
    print(f"case trained HfO2 NEP {dtype}: {len(frames)} frames x {frames[0].natoms} atoms", flush=True)
    want = on.nep_evaluate(frames, w)
    cfg = w.cfg
    gp = mn.pack_frames(frames, cfg.rcut_radial, dev)
    nb = mn.neighbors(gp, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes)
    model = mn.NEPTorch(w, dtype, dev)
    Etot, Ei, F, V = model(nb["NN_r"], nb["NL_r"], nb["Ri_r"], nb["NN_a"], nb["NL_a"], nb["Ri_a"], gp["num_atoms"],
                           gp["atom_type_map"], create_graph=False)
    assert _rel(Etot.detach().cpu().double().reshape(-1), want["E"]) < tol
    assert _rel(F.detach().cpu().double(), want["F"]) < 100 * tol
    assert _rel(V.detach().cpu().double(), want["V"]) < 100 * tol
    Fp = F.detach().cpu().double().numpy().reshape(len(frames), -1)
    assert float(np.sqrt(((Fp - d["forces"]) ** 2).mean())) < 0.1


def test_model_float32(dev):
    frames = _frames()
    cfg = NEPConfig(ntypes=4, l_max=[4, 2, 0])
    w = NEPWeights.random(cfg, seed=14)
    want = on.nep_evaluate(frames, w)
    gp = mn.pack_frames(frames, cfg.rcut_radial, dev)
    nb = mn.neighbors(gp, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes)
    model = mn.NEPTorch(w, torch.float32, dev)
    Etot, Ei, F, V = model(nb["NN_r"], nb["NL_r"], nb["Ri_r"], nb["NN_a"], nb["NL_a"], nb["Ri_a"], gp["num_atoms"],
                           gp["atom_type_map"], create_graph=False)
    assert _rel(Etot.detach().cpu().reshape(-1), want["E"]) < 1e-5
    f_rmse = float(np.sqrt(np.mean((F.detach().cpu().double().numpy() - want["F"]) ** 2)))
    assert f_rmse < 1e-4, f_rmse


def test_force_loss_training_gradient(dev):
    """d(sum F^2 + E^2)/d(c2, c3, fitting params) through the double-backward chain vs the oracle."""
    frames = [st.zincblende_sic((2, 2, 2), seed=4)]
    cfg = NEPConfig(ntypes=2, l_max=[4, 2, 1])
    w = NEPWeights.random(cfg, seed=15)
    pk = on.pack_frames(frames, cfg.rcut_radial)
    onb = on.neighbor(pk, 6.0, 4.0, 120, 40)
    am = pk["atom_type_map"]
    # oracle: torch FP64 CPU with parameters as leaves
    c2 = torch.tensor(w.c2, requires_grad=True)
    c3 = torch.tensor(w.c3, requires_grad=True)
    Ri_r = torch.as_tensor(onb["Ri_r"]).clone().requires_grad_()
    Ri_a = torch.as_tensor(onb["Ri_a"]).clone().requires_grad_()
    q = torch.cat([on.radial_feats(c2, Ri_r, onb["NL_r"], am, 6.0),
                   on.angular_feats(c3, Ri_a, onb["NL_a"], am, 4.0, 4, 2, 1)], -1)
    fit = [[(torch.tensor(W, requires_grad=True), torch.tensor(b, requires_grad=True),
             None if r is None else torch.tensor(r, requires_grad=True)) for (W, b, r) in layers] for layers in w.fit]
    x = q * torch.as_tensor(w.q_scaler)
    amt = torch.as_tensor(am)
    Ei = torch.zeros(q.shape[0], dtype=torch.float64)
    for t in range(cfg.ntypes):
        m = amt == t
        Ei = Ei.index_put((m,), on.fit_forward(fit[t], x[m]).reshape(-1))
    E = Ei.sum()
    gr, ga = torch.autograd.grad(E, [Ri_r, Ri_a], create_graph=True)
    loss_terms = []
    for Ri, gg, NL in ((Ri_r, gr, onb["NL_r"]), (Ri_a, ga, onb["NL_a"])):
        NLt = torch.as_tensor(NL)
        valid = (NLt >= 0).unsqueeze(-1)
        r = torch.where(valid[..., 0], Ri[..., 0].detach(), torch.ones_like(Ri[..., 0]))
        pair = (gg[..., 0:1] * Ri[..., 1:4].detach() / r.unsqueeze(-1) + gg[..., 1:4]) * valid
        F = pair.sum(1).index_add(0, NLt.clamp(min=0).reshape(-1), -pair.reshape(-1, 3))
        loss_terms.append(F)
    F_or = loss_terms[0] + loss_terms[1]
    loss = (F_or ** 2).sum() + E ** 2
    params_or = [c2, c3] + [p for layers in fit for (W, b, r) in layers for p in (W, b, r) if p is not None]
    want = torch.autograd.grad(loss, params_or)
    # HIP path
    gp = mn.pack_frames(frames, cfg.rcut_radial, dev)
    nb = mn.neighbors(gp, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes, 120, 40)
    model = mn.NEPTorch(w, torch.float64, dev)
    Etot, _, Fg, _ = model(nb["NN_r"], nb["NL_r"], nb["Ri_r"], nb["NN_a"], nb["NL_a"], nb["Ri_a"], gp["num_atoms"],
                           gp["atom_type_map"], create_graph=True)
    assert _rel(Fg.detach().cpu(), F_or.detach()) < 1e-10
    lg = (Fg ** 2).sum() + (Etot ** 2).sum()
    params = [model.c2, model.c3]
    for net in model.fit:
        for l in range(len(net.W)):
            params += [net.W[l], net.b[l]] + ([net.rdt[l]] if net.rdt[l] is not None else [])
    got = torch.autograd.grad(lg, params)
    for a, b in zip(got, want):
        assert _rel(a.detach().cpu(), b) < 1e-9


def test_graph_replays_match_eager(dev):
    """NEPGraph (neighbor search + features + net + forces + virial captured as one HIP graph, no host
    synchronisation inside): three replays, positions moved in place before each and eager work in
    between, equal the checked eager path at the same positions (energies and forces to 1e-12
    relative; the frame virial sums with atomics in either order, 1e-12 relative)."""
    f = st.zincblende_sic((3, 3, 3), a=4.36, sigma=0.08, seed=5)
    cfg = NEPConfig(ntypes=2, l_max=[4, 2, 0])
    w = NEPWeights.random(cfg, seed=7)
    gp = mn.pack_frames([f], cfg.rcut_radial, dev)
    NNr, NNa = nep_ops.calculate_maxneigh(gp["num_atoms"], gp["box"], gp["box_orig"], gp["num_cell"],
                                          gp["position"], cfg.rcut_radial, cfg.rcut_angular, 2, gp["atom_type_map"],
                                          False)
    maxr, maxa = int(NNr.max().item()) + 8, int(NNa.max().item()) + 8
    model = mn.NEPTorch(w, torch.float64, dev)
    g = mn.NEPGraph(model, gp, 2, maxr, maxa)
    rng = np.random.default_rng(0)
    pos0 = gp["position"].clone()
    for rep in range(3):
        print(f"case NEP graph replay {rep}", flush=True)
        gp["position"].copy_(pos0 + torch.as_tensor(rng.normal(0.0, 0.01, tuple(pos0.shape)), device=dev))
        E, Ei, F, V = [t.detach().clone() for t in g.replay()]
        nb = mn.neighbors(gp, cfg.rcut_radial, cfg.rcut_angular, 2, maxr, maxa)
        E0, Ei0, F0, V0 = model(nb["NN_r"], nb["NL_r"], nb["Ri_r"], nb["NN_a"], nb["NL_a"], nb["Ri_a"],
                                gp["num_atoms"], gp["atom_type_map"], create_graph=False)
        torch.cuda.synchronize()
        for a, b in ((E, E0), (Ei, Ei0), (F, F0), (V, V0)):
            b = b.detach()
            assert torch.allclose(a, b, rtol=1e-12, atol=1e-12 * float(b.abs().max())), (rep, (a - b).abs().max())
    g.check()
