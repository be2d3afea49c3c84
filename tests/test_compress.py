"""DP compress on the CPU: the oracle's polynomial evaluation against numpy.polyval, and the table
builder (matpl_amd.compress.build_table) against the Hermite-interpolation properties that fix it:
exact reproduction of cubic (order 3) / quintic (order 5) functions, values and derivatives of the
net at every knot, and the interpolation error bound on the real embedding net."""
import math

import numpy as np
import pytest
import torch

import oracle.compress as oc
from matpl_amd.compress import build_table


def test_oracle_horner_vs_polyval():
    rng = np.random.default_rng(0)
    for K in (4, 6):
        coef = rng.normal(size=(7, 3, K))
        f2 = rng.normal(size=7)
        G = oc.horner(f2, coef)
        dG = oc.horner_d(f2, coef)
        for s in range(7):
            for i in range(3):
                assert abs(G[s, i] - np.polyval(coef[s, i], f2[s])) < 1e-12
                assert abs(dG[s, i] - np.polyval(np.polyder(coef[s, i]), f2[s])) < 1e-12


def _poly_nets(deg, T=2, E=3, seed=1):
    rng = np.random.default_rng(seed)
    cs = [torch.as_tensor(rng.normal(size=(E, deg + 1))) for _ in range(T * T)]

    def make(c):
        def net(s):                               # s [n, 1] -> [n, E], polynomial of degree deg
            out = torch.zeros(s.shape[0], c.shape[0], dtype=torch.float64)
            for q in range(c.shape[1]):
                out = out * s + c[:, q]
            return out
        return net
    return [make(c) for c in cs], cs


def _eval(ct, s, pair):
    f2, coef = ct.lookup(torch.as_tensor(s), pair)
    return oc.horner(f2.numpy(), coef.numpy())


@pytest.mark.parametrize("order,deg", [(3, 3), (5, 5)])
def test_table_reproduces_polynomials_exactly(order, deg):
    nets, cs = _poly_nets(deg)
    davg = np.array([[0.3, 0, 0, 0], [0.2, 0, 0, 0]])
    dstd = np.array([[0.1, 1, 1, 1], [0.15, 1, 1, 1]])
    ct = build_table(nets, davg, dstd, s_max_raw=1.2, dx=0.05, order=order)
    rng = np.random.default_rng(2)
    s = np.concatenate([rng.uniform(ct.x0, 10 * ct.x_fine_end, 500), ct.knots()[:-1]])
    for p, c in enumerate(cs):
        want = np.stack([np.polyval(c[i].numpy(), s) for i in range(c.shape[0])], -1)
        got = _eval(ct, s, p)
        np.testing.assert_allclose(got, want, rtol=1e-8, atol=1e-8 * np.abs(want).max())


def _dp_emb_nets(seed=3):
    from matpl_amd.dp import _Net
    from matpl_amd.dpnet import DPConfig, DPWeights
    cfg = DPConfig(ntypes=2)
    w = DPWeights.random(cfg, seed=seed)
    return [_Net(w.emb[p], 0, torch.float64, "cpu") for p in range(4)]


@pytest.mark.parametrize("order", [3, 5])
def test_table_matches_net_at_knots_and_between(order):
    nets = _dp_emb_nets()
    davg = np.array([[0.05, 0, 0, 0], [0.04, 0, 0, 0]])
    dstd = np.array([[0.08, 1, 1, 1], [0.07, 1, 1, 1]])
    dx = 0.01
    ct = build_table(nets, davg, dstd, s_max_raw=0.6, dx=dx, order=order)
    kn = ct.knots()
    mid = 0.5 * (kn[1:] + kn[:-1])
    for p, net in enumerate(nets):
        xs = torch.as_tensor(kn[:-1]).unsqueeze(-1).requires_grad_()
        y = net(xs)
        # at the left knot of every interval f2 = 0: value and slope are the net's
        rows = ct.table[p].numpy()
        np.testing.assert_allclose(rows[:, :, -1], y.detach().numpy(), rtol=0, atol=1e-12)
        dy = torch.stack([torch.autograd.grad(y[:, i].sum(), xs, retain_graph=True)[0][:, 0]
                          for i in range(y.shape[1])], -1)
        np.testing.assert_allclose(rows[:, :, -2], dy.numpy(), rtol=0, atol=1e-10)
        # at the right knot the polynomial reaches the next knot's value
        h = np.diff(kn)
        right = oc.horner(h, rows)
        yr = net(torch.as_tensor(kn[1:]).unsqueeze(-1)).detach().numpy()
        np.testing.assert_allclose(right, yr, rtol=0, atol=1e-10)
        # midpoints of the fine part: Hermite error O(dx^4) (cubic) / O(dx^6) (quintic)
        fine = mid < ct.x_fine_end
        ym = net(torch.as_tensor(mid[fine]).unsqueeze(-1)).detach().numpy()
        err = np.abs(_eval(ct, mid[fine], p) - ym).max()
        assert err < (1e-7 if order == 3 else 1e-10), err
