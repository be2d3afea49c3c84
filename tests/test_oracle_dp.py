"""Pins the FP64 DP oracle against what the mathematics fixes.

The reference ships no golden vectors for this path, so each oracle routine
is checked against closed forms, invariants and finite differences chosen so
that a dropped term, a wrong sign or a transposed index fails:
  * neighbor search: fcc shell counts (78 within 6 A for a = 3.615 A),
    pair symmetry j in N(i) <=> i in N(j) with dR_ji = -dR_ij, per-type
    buckets, and equality with an independent numpy search over images;
  * env mat: central finite differences of Ri wrt dR, continuity of s(r);
  * energy: forces = -dE/dx by central differences, virial = -dE/d(strain)
    by central differences, sum of forces = 0, rotation invariance of E,
    permutation covariance of F;
  * operator contract: calculateForce/Virial and their backward equal a
    dense numpy scatter and the transpose (adjoint) identity.
"""
import numpy as np
import pytest

import oracle
from matpl_amd import structures as st
from matpl_amd.dpnet import DPConfig, DPWeights


def _numpy_neighbors(frame, rc):
    """Independent O(N^2 * images) search with explicit image enumeration."""
    L = frame.lattice
    frac = np.mod(frame.frac() + 10.0, 1.0)
    n = frame.natoms
    imgs = np.array([[a, b, c] for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3)], float)
    out = []
    for i in range(n):
        d = (frac[None, :, :] - frac[i][None, None, :]) + imgs[:, None, :]
        cart = d @ L
        r2 = (cart ** 2).sum(-1)
        ok = (r2 < rc * rc) & (r2 > 1e-8)
        out.append(sorted(np.nonzero(ok)[1].tolist()))
    return out


def test_fcc_shell_count():
    f = st.fcc_cu(sigma=0.0)
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 1, 100, 6.0)
    cnt = (lst > 0).sum(1)
    assert np.all(cnt == 78)            # 12 + 6 + 24 + 12 + 24 shells below 6 A
    r = np.sqrt((dR ** 2).sum(-1))[lst > 0]
    shells = np.unique(np.round(r, 6))
    a = 3.615
    np.testing.assert_allclose(shells, a * np.sqrt([0.5, 1.0, 1.5, 2.0, 2.5]), atol=1e-6)
    for rs, c in zip(a * np.sqrt([0.5, 1.0, 1.5, 2.0, 2.5]), [12, 6, 24, 12, 24]):
        assert np.all((np.abs(np.sqrt((dR ** 2).sum(-1)) - rs) < 1e-9).sum(1) == c)


@pytest.mark.parametrize("maker,rc", [(lambda: st.fcc_cu(), 6.0), (lambda: st.zincblende_sic((2, 2, 2)), 6.0),
                                      (lambda: st.fluorite_hfo2((3, 3, 3)), 6.0)])
def test_neighbor_symmetry_and_independent_search(maker, rc):
    f = maker()
    T = len(f.elements)
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, T, 200, rc)
    n = f.natoms
    m = 200
    # per-type buckets hold only neighbors of that type
    for t in range(T):
        js = lst[:, t * m:(t + 1) * m]
        assert np.all(f.types[js[js > 0] - 1] == t)
    # multiset of (i, j, dR) is antisymmetric
    pairs = {}
    for i in range(n):
        for s in np.nonzero(lst[i])[0]:
            key = (i, lst[i, s] - 1, tuple(np.round(dR[i, s], 9)))
            pairs[key] = pairs.get(key, 0) + 1
    for (i, j, d), c in pairs.items():
        assert pairs.get((j, i, tuple(-np.asarray(d))), 0) == c
    # same membership as an independent search (multiplicity counts images)
    ref = _numpy_neighbors(f, rc)
    for i in range(n):
        got = sorted((lst[i][lst[i] > 0] - 1).tolist())
        assert got == ref[i]


def test_small_box_images_and_order():
    # 2x2x2 fcc: L = 7.23 < 2 rc, so the same j appears under several images
    f = st.fcc_cu((2, 2, 2), sigma=0.05, seed=7)
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 1, 200, 6.0)
    row = lst[0][lst[0] > 0]
    assert len(row) > len(set(row.tolist()))           # repeated j
    assert np.all(np.diff(row) >= 0)                    # (j, image) order: j non-decreasing


def test_overflow_raises():
    f = st.fcc_cu()
    with pytest.raises(OverflowError):
        oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 1, 50, 6.0)


def test_env_mat_finite_difference():
    rng = np.random.default_rng(3)
    n, S = 3, 8
    dR = rng.uniform(-4, 4, size=(n, S, 3))
    dR[:, -1] = 0.0                                      # one empty slot per atom
    lst = np.ones((n, S), np.int32)
    lst[:, -1] = 0
    davg = np.array([[0.1, 0, 0, 0]])
    dstd = np.array([[0.3, 0.2, 0.2, 0.2]])
    tm = np.zeros(n, np.int32)
    rmin, rmax = 2.0, 6.0
    Ri, Ri_d = oracle.dp_env_mat(oracle.dp_image_dr(lst, dR), tm, davg, dstd, rmin, rmax)
    h = 1e-6
    for a in range(3):
        p = dR.copy(); p[..., a] += h
        q = dR.copy(); q[..., a] -= h
        Rp, _ = oracle.dp_env_mat(oracle.dp_image_dr(lst, p), tm, davg, dstd, rmin, rmax)
        Rq, _ = oracle.dp_env_mat(oracle.dp_image_dr(lst, q), tm, davg, dstd, rmin, rmax)
        fd = (Rp - Rq) / (2 * h)                         # d Ri / d dR_a
        # Ri_d = -dfeat/dstd = -d Ri_norm / d dR
        np.testing.assert_allclose(-Ri_d[..., a], fd, atol=1e-7, rtol=1e-6)
    # empty slots: Ri = -davg/dstd, Ri_d = 0
    np.testing.assert_allclose(Ri[:, -1], np.broadcast_to(-davg[0] / dstd[0], (n, 4)))
    assert np.all(Ri_d[:, -1] == 0)


def test_switch_continuity():
    rmin, rmax = 5.5, 6.0
    lst = np.ones((1, 4), np.int32)
    dR = np.zeros((1, 4, 3))
    dR[0, :, 0] = [rmin - 1e-9, rmin + 1e-9, rmax - 1e-9, rmax + 1e-9]
    Ri, _ = oracle.dp_env_mat(oracle.dp_image_dr(lst, dR), np.zeros(1, np.int32), np.zeros((1, 4)),
                              np.ones((1, 4)), rmin, rmax)
    assert abs(Ri[0, 0, 0] - Ri[0, 1, 0]) < 1e-7
    assert abs(Ri[0, 2, 0]) < 1e-7 and Ri[0, 3, 0] == 0.0


def _cu_model(seed=11):
    f = st.fcc_cu()
    cfg = DPConfig(ntypes=1, m=100)
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 1, 100, 6.0)
    davg, dstd = oracle.dp_calc_stat(lst, dR, f.types, 1, cfg.rcut_smooth, cfg.rcut)
    return f, DPWeights.random(cfg, seed=seed, davg=davg, dstd=dstd)


def _hfo2_model(seed=12):
    f = st.fluorite_hfo2((3, 3, 3), sigma=0.1, seed=5)
    cfg = DPConfig(ntypes=2, m=120)
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 2, 120, 6.0)
    davg, dstd = oracle.dp_calc_stat(lst, dR, f.types, 2, cfg.rcut_smooth, cfg.rcut)
    return f, DPWeights.random(cfg, seed=seed, davg=davg, dstd=dstd)


def _energy(frame, w):
    return oracle.dp_evaluate(frame, w)["E"]


@pytest.mark.parametrize("model", [_cu_model, _hfo2_model])
def test_forces_are_minus_energy_gradient(model):
    f, w = model()
    r = oracle.dp_evaluate(f, w)
    h = 1e-4
    rng = np.random.default_rng(0)
    for a in rng.choice(f.natoms, 4, replace=False):
        for c in range(3):
            p = st.Frame(f.lattice, f.pos.copy(), f.types, f.elements); p.pos[a, c] += h
            q = st.Frame(f.lattice, f.pos.copy(), f.types, f.elements); q.pos[a, c] -= h
            fd = -(_energy(p, w) - _energy(q, w)) / (2 * h)
            assert abs(fd - r["F"][a, c]) < 1e-6 * max(1.0, abs(fd))
    np.testing.assert_allclose(r["F"].sum(0), 0.0, atol=1e-10)


@pytest.mark.parametrize("model", [_cu_model, _hfo2_model])
def test_virial_is_minus_strain_derivative(model):
    f, w = model()
    r = oracle.dp_evaluate(f, w)
    eps = 1e-5
    for a in range(3):
        for b in range(3):
            E = []
            for sgn in (1, -1):
                M = np.eye(3); M[a, b] += sgn * eps
                E.append(_energy(st.Frame(f.lattice @ M, f.pos @ M, f.types, f.elements), w))
            fd = -(E[0] - E[1]) / (2 * eps)
            # x' = x (I + eps e_a e_b^T): W_ab = sum dR_a f_b; the CPU branch fills the
            # lower triangle from the upper one, exact because W is symmetric
            assert abs(fd - r["V"][3 * a + b]) < 2e-5 * max(1.0, abs(fd))


def test_rotation_and_permutation():
    f, w = _cu_model()
    r = oracle.dp_evaluate(f, w)
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    fr = st.Frame(f.lattice @ R.T, f.pos @ R.T, f.types, f.elements)
    rr = oracle.dp_evaluate(fr, w)
    assert abs(rr["E"] - r["E"]) < 1e-9 * abs(r["E"])
    np.testing.assert_allclose(rr["F"], r["F"] @ R.T, atol=1e-9)
    perm = np.random.default_rng(1).permutation(f.natoms)
    fp = st.Frame(f.lattice, f.pos[perm], f.types[perm], f.elements)
    rp = oracle.dp_evaluate(fp, w)
    assert abs(rp["E"] - r["E"]) < 1e-9 * abs(r["E"])
    np.testing.assert_allclose(rp["F"], r["F"][perm], atol=1e-9)


def test_padding_term_depends_on_m():
    # empty slots enter A through G(-davg/dstd) (dp_dp.py:293-297): E changes with m
    f, w = _cu_model()
    e100 = oracle.dp_evaluate(f, w)["E"]
    cfg2 = DPConfig(ntypes=1, m=120)
    w2 = DPWeights(cfg2, w.emb, w.fit, w.davg, w.dstd)
    e120 = oracle.dp_evaluate(f, w2)["E"]
    assert abs(e100 - e120) > 1e-6


def test_calc_stat_closed_form():
    f = st.fcc_cu(sigma=0.0)
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 1, 100, 6.0)
    davg, dstd = oracle.dp_calc_stat(lst, dR, f.types, 1, 5.5, 6.0)
    # perfect fcc: each atom has the same 78 neighbors; s = vv/r over the shells
    a = 3.615
    shells = a * np.sqrt([0.5, 1.0, 1.5, 2.0, 2.5])
    counts = [12, 6, 24, 12, 24]

    def s_of(r):
        if r < 5.5:
            return 1 / r
        u = (r - 5.5) / 0.5
        return (u ** 3 * (-6 * u * u + 15 * u - 10) + 1) / r
    s = np.array([s_of(r) for r in shells])
    mean = (counts * s).sum() / 100.0
    std = np.sqrt((counts * s * s).sum() / 100.0 - mean ** 2)
    assert abs(davg[0, 0] - mean) < 1e-12
    assert abs(dstd[0, 0] - std) < 1e-12


def test_operator_contract_dense_scatter():
    rng = np.random.default_rng(5)
    B, N, S, ng = 2, 5, 6, 2
    nl = rng.integers(-1, N + ng, size=(B, N, S)).astype(np.int32)
    dE = rng.normal(size=(B, N, S * 4))
    Rid = rng.normal(size=(B, N, S * 4, 3))
    Rij = rng.normal(size=(B, N, S, 3))
    Fc = rng.normal(size=(B, N, 3))
    F = oracle.dp_force_op(nl, dE, Rid, Fc, ng)
    want = np.zeros((B, N + ng, 3))
    want[:, :N] = Fc
    per = np.einsum("bnsw,bnswa->bnsa", dE.reshape(B, N, S, 4), Rid.reshape(B, N, S, 4, 3))
    for b in range(B):
        for i in range(N):
            for s in range(S):
                if nl[b, i, s] >= 0:
                    want[b, nl[b, i, s]] += per[b, i, s]
    np.testing.assert_allclose(F, want, atol=1e-12)
    V, AV = oracle.dp_virial_op(nl, dE, Rij, Rid, ng)
    wantA = np.zeros((B, N + ng, 9))
    for b in range(B):
        for i in range(N):
            for s in range(S):
                if nl[b, i, s] >= 0:
                    wantA[b, nl[b, i, s]] += np.outer(per[b, i, s], Rij[b, i, s]).ravel()
    np.testing.assert_allclose(AV, wantA, atol=1e-12)
    np.testing.assert_allclose(V, wantA.sum(1), atol=1e-12)
    # backward ops are the adjoints of the forward ops (nghost = 0 so j mod N is the identity)
    nl0 = np.where(nl >= N, -1, nl).astype(np.int32)
    gF = rng.normal(size=(B, N, 3))
    g = oracle.dp_force_grad_op(nl0, Rid, gF, 0)
    lin = lambda x: oracle.dp_force_op(nl0, x, Rid, -np.einsum("bnsw,bnswa->bna", x.reshape(B, N, S, 4),
                                                                 Rid.reshape(B, N, S, 4, 3)), 0)
    x = rng.normal(size=dE.shape)
    assert abs((lin(x) * gF).sum() - (g * x).sum()) < 1e-9
    gV = rng.normal(size=(B, 9))
    gv = oracle.dp_virial_grad_op(nl0, Rij, Rid, gV)
    assert abs((oracle.dp_virial_op(nl0, x, Rij, Rid, 0)[0] * gV).sum() - (gv * x).sum()) < 1e-9


def _drop_type(w, t_drop):
    """The (T-1)-type model that keeps every type but t_drop: its nets, davg/dstd rows."""
    T = w.cfg.ntypes
    keep = [t for t in range(T) if t != t_drop]
    cfg = DPConfig(ntypes=T - 1, m=w.cfg.m, rcut=w.cfg.rcut, rcut_smooth=w.cfg.rcut_smooth)
    emb = [w.emb[a * T + b] for a in keep for b in keep]
    return DPWeights(cfg, emb, [w.fit[t] for t in keep], w.davg[keep], w.dstd[keep]), keep


@pytest.mark.parametrize("t_drop", [0, 1, 2])
def test_absent_type_equals_reduced_model(t_drop):
    """A frame without type t, evaluated with t left out of its type list (dp_dp.py:150-165, 197,
    413), equals the (T-1)-type model with t's nets, bucket and statistics removed; with t listed
    (every model type) it does not, because the empty t bucket then adds its padding term and
    type_nums counts t."""
    base = st.fluorite_hfo2((2, 2, 2), seed=41)
    keep2 = [t for t in range(3) if t != t_drop]
    f3 = st.Frame(base.lattice, base.pos, np.asarray(keep2, np.int32)[base.types], ["A", "B", "C"])
    cfg = DPConfig(ntypes=3, m=100)
    rng = np.random.default_rng(7)
    davg = np.c_[rng.uniform(0.05, 0.2, 3), np.zeros((3, 3))]
    dstd = np.c_[rng.uniform(0.05, 0.2, 3), rng.uniform(0.05, 0.2, (3, 3))]
    w3 = DPWeights.random(cfg, seed=12, davg=davg, dstd=dstd)
    tp = np.ones(3, np.int32)
    tp[t_drop] = 0
    r3 = oracle.dp_evaluate(f3, w3, type_present=tp)
    w2, keep = _drop_type(w3, t_drop)
    f2 = st.Frame(base.lattice, base.pos, base.types.copy(), ["A", "B"])
    r2 = oracle.dp_evaluate(f2, w2)
    assert abs(r3["E"] - r2["E"]) <= 1e-12 * abs(r2["E"])
    np.testing.assert_allclose(r3["F"], r2["F"], atol=1e-12)
    np.testing.assert_allclose(r3["V"], r2["V"], atol=1e-11 * np.abs(r2["V"]).max())
    full = oracle.dp_evaluate(f3, w3)
    assert abs(full["E"] - r2["E"]) > 1e-6 * abs(r2["E"])
