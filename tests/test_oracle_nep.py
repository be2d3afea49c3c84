"""Pins the NEP oracle (oracle/nep_oracle.c + oracle/nep.py).

Against the reference's own CPU code compiled from source into oracle/_ref
(oracle/ref.mk; skipped where it is not built): neighbor lists bit-exact,
per-type neighbor counts exact, descriptor features to 1e-12 relative,
and the helper functions (fc, Chebyshev f_n, s_lm accumulation, q with the
C3B/C4B/C5B constants). Against the mathematics: T_k(x) = cos(k acos x),
rotation invariance of the features, forces = -dE/dx and virial =
-dE/d(strain) by central differences, sum of forces = 0.
"""
import math

import numpy as np
import pytest
import torch

import oracle.nep as on
import oracle.ref as ref
from matpl_amd import structures as st
from matpl_amd.nepnet import NEPConfig, NEPWeights

needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (needs the reference tree)")


def _frames():
    return [st.zincblende_sic((3, 3, 3), seed=3), st.zincblende_sic((2, 2, 2), seed=4),
            st.agausic_frames(nframes=2, nmin=32, nmax=60, seed=5)[0]]


@needs_ref
def test_neighbor_matches_reference_bit_exact():
    frames = _frames()
    pk = on.pack_frames(frames, 6.0)
    a = on.neighbor(pk, 6.0, 4.0, 400, 80)
    b = ref.neighbor(pk, 6.0, 4.0, 400, 80)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    for with_type in (False, True):
        r1, a1 = on.maxneigh(pk, 6.0, 4.0, 4, with_type)
        r2, a2 = ref.maxneigh(pk, 6.0, 4.0, 4, with_type)
        assert np.array_equal(r1, r2) and np.array_equal(a1, a2)


@needs_ref
@pytest.mark.parametrize("lmax", [(4, 2, 0), (4, 2, 1), (4, 0, 0)])
def test_descriptor_matches_reference(lmax):
    f = st.zincblende_sic((3, 3, 3), seed=3)
    cfg = NEPConfig(ntypes=2, l_max=list(lmax))
    w = NEPWeights.random(cfg, seed=7)
    pk = on.pack_frames([f], cfg.rcut_radial)
    # the reference descriptor takes the radial list and [x, y, z], filtering r <= rc_angular itself
    b = ref.neighbor(pk, 6.0, 6.0, 120, 120, with_rij=False)
    want = ref.descriptor(w.c2, w.c3, b["Ri_r"], b["NL_r"], pk["atom_type_map"], 6.0, 4.0, *lmax)
    a = on.neighbor(pk, 6.0, 4.0, 120, 60)
    q = torch.cat([on.radial_feats(w.c2, torch.as_tensor(a["Ri_r"]), a["NL_r"], pk["atom_type_map"], 6.0),
                   on.angular_feats(w.c3, torch.as_tensor(a["Ri_a"]), a["NL_a"], pk["atom_type_map"], 4.0, *lmax)],
                  -1).numpy()
    np.testing.assert_allclose(q, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@needs_ref
def test_helpers_match_reference():
    rng = np.random.default_rng(0)
    for _ in range(20):
        rc = 5.0
        d = rng.uniform(0.5, 5.5)
        fc = ref.find_fc(rc, d)
        assert abs(fc - (0.5 * math.cos(math.pi * d / rc) + 0.5 if d < rc else 0.0)) < 1e-15
        fn = ref.find_fn(12, rc, d, fc)
        mine = on.fk(torch.tensor([d], dtype=torch.float64), 13, rc)[0].numpy()
        np.testing.assert_allclose(mine, fn, rtol=1e-13, atol=1e-15)
        xyz = rng.normal(size=3)
        r = float(np.linalg.norm(xyz))
        s = ref.accumulate_s(r, *xyz, 1.0)
        b = on.blm(torch.tensor([r], dtype=torch.float64), torch.tensor(np.asarray([xyz])))[0].numpy()
        np.testing.assert_allclose(b, s, rtol=1e-13, atol=1e-14)
    na = 3
    sv = rng.normal(size=(na, 24))
    for kind, l4, l5 in (("find_q", 0, 0), ("find_q_with_4body", 2, 0), ("find_q_with_5body", 2, 1)):
        q = np.zeros(6 * na)
        for n in range(na):
            q += ref.find_q(kind, na, n, sv[n])
        mine = on.q_from_s(torch.as_tensor(sv)[None], 4, l4, l5)[0].numpy()
        np.testing.assert_allclose(mine, q[:mine.shape[0]], rtol=1e-13, atol=1e-13)


def test_chebyshev_closed_form():
    r = torch.linspace(0.1, 5.9, 50, dtype=torch.float64)
    rc = 6.0
    f = on.fk(r, 13, rc)
    x = 2 * (r / rc - 1) ** 2 - 1
    fc = 0.5 * torch.cos(math.pi * r / rc) + 0.5
    for k in range(13):
        want = (torch.cos(k * torch.arccos(x)) + 1) * 0.5 * fc
        assert torch.allclose(f[:, k], want, atol=1e-12)


def test_features_rotation_invariant():
    f = st.zincblende_sic((3, 3, 3), seed=3)
    cfg = NEPConfig(ntypes=2, l_max=[4, 2, 1])
    w = NEPWeights.random(cfg, seed=7)
    r0 = on.nep_evaluate([f], w)
    th = 0.7
    R = np.array([[1, 0, 0], [0, np.cos(th), -np.sin(th)], [0, np.sin(th), np.cos(th)]]) @ \
        np.array([[np.cos(0.3), 0, np.sin(0.3)], [0, 1, 0], [-np.sin(0.3), 0, np.cos(0.3)]])
    r1 = on.nep_evaluate([st.Frame(f.lattice @ R.T, f.pos @ R.T, f.types, f.elements)], w)
    np.testing.assert_allclose(r1["q"], r0["q"], rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(r1["F"], r0["F"] @ R.T, atol=1e-9)


@pytest.mark.parametrize("idx", [0, 1])
def test_forces_and_virial_finite_difference(idx):
    frames = [st.zincblende_sic((2, 2, 2), seed=4), st.agausic_frames(nframes=1, nmin=40, nmax=40, seed=8)[0]]
    f = frames[idx]
    T = len(f.elements)
    w = NEPWeights.random(NEPConfig(ntypes=T, l_max=[4, 2, 1]), seed=9)
    r = on.nep_evaluate([f], w)
    np.testing.assert_allclose(r["F"].sum(0), 0.0, atol=1e-10)
    h = 1e-4
    for (i, c) in [(0, 0), (5, 1), (f.natoms - 1, 2)]:
        p = st.Frame(f.lattice, f.pos.copy(), f.types, f.elements); p.pos[i, c] += h
        m = st.Frame(f.lattice, f.pos.copy(), f.types, f.elements); m.pos[i, c] -= h
        fd = -(on.nep_evaluate([p], w)["E"][0] - on.nep_evaluate([m], w)["E"][0]) / (2 * h)
        assert abs(fd - r["F"][i, c]) < 1e-6 * max(1.0, abs(fd))
    eps = 1e-5
    for (a, b) in [(0, 0), (1, 2), (2, 1)]:
        E = []
        for sg in (1, -1):
            M = np.eye(3); M[a, b] += sg * eps
            E.append(on.nep_evaluate([st.Frame(f.lattice @ M, f.pos @ M, f.types, f.elements)], w)["E"][0])
        fd = -(E[0] - E[1]) / (2 * eps)
        assert abs(fd - r["V"][0][3 * a + b]) < 2e-5 * max(1.0, abs(fd))


def test_differentiable_efv_equals_pinned_evaluate():
    """oracle.nep.nep_efv_graph (the torch form used for loss gradients) gives the energies, forces and
    virials of nep_evaluate, which the position and strain differences above pin."""
    frames = [st.zincblende_sic((2, 2, 2), seed=4), st.agausic_frames(nframes=2, nmin=30, nmax=50, seed=5)[0]]
    cfg = NEPConfig(ntypes=4, l_max=[4, 2, 1])
    w = NEPWeights.random(cfg, seed=3)
    want = on.nep_evaluate(frames, w)
    fit = [[(torch.tensor(W), torch.tensor(b), None if r is None else torch.tensor(r)) for W, b, r in L]
           for L in w.fit]
    E, F, V = on.nep_efv_graph(frames, cfg, torch.tensor(w.c2), torch.tensor(w.c3), fit, w.q_scaler)
    np.testing.assert_allclose(E.detach().numpy(), want["E"], rtol=1e-13, atol=0)
    np.testing.assert_allclose(F.detach().numpy(), want["F"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(V.detach().numpy(), want["V"], rtol=0, atol=1e-12 * np.abs(want["V"]).max())
