"""Host-side NEP packing (matpl_amd.nep.pack_frames / expand_box) against the oracle's C restatement of
expand_box, bit for bit, including skewed and small cells that need several replicas."""
import numpy as np
import torch

import oracle.nep as on
from matpl_amd import nep as mn
from matpl_amd import structures as st


def test_pack_frames_matches_oracle():
    frames = [st.zincblende_sic((2, 2, 2), seed=4), st.zincblende_sic((1, 1, 1), seed=1),
              st.agausic_frames(nframes=2, nmin=30, nmax=50, seed=5)[0], st.fcc_cu((2, 2, 2), seed=2)]
    rng = np.random.default_rng(0)
    skew = st.Frame(frames[0].lattice + rng.normal(scale=0.4, size=(3, 3)), frames[0].pos, frames[0].types,
                    frames[0].elements)
    frames.append(skew)
    for rc in (6.0, 4.0, 8.5):
        want = on.pack_frames(frames, rc)
        got = mn.pack_frames(frames, rc, "cpu")
        for k in want:
            assert np.array_equal(got[k].numpy().reshape(-1), np.asarray(want[k]).reshape(-1)), (k, rc)


def test_tall_products_match_plain_autograd():
    """The fitting layers' split-K product: values, gradients and second derivatives equal the plain
    x @ W ones (row counts below, at and above the split threshold, with padding)."""
    from matpl_amd.nep import _SPLIT_ROWS, _TallMM
    torch.manual_seed(0)
    for K in (100, 2 * _SPLIT_ROWS + 1, 3 * _SPLIT_ROWS + 7):
        x = torch.randn(K, 7, dtype=torch.float64, requires_grad=True)
        W = torch.randn(7, 5, dtype=torch.float64, requires_grad=True)
        outs = []
        for f in (lambda a, b: _TallMM.apply(a, b), lambda a, b: a @ b):
            xx, WW = x.detach().clone().requires_grad_(), W.detach().clone().requires_grad_()
            y = torch.tanh(f(xx, WW))
            (gx,) = torch.autograd.grad(y.pow(2).sum(), xx, create_graph=True)
            gW = torch.autograd.grad((gx ** 2).sum(), WW)[0]
            outs.append((y.detach(), gx.detach(), gW))
        for a, b in zip(*outs):
            assert torch.allclose(a, b, rtol=1e-12, atol=1e-10)
    xs = torch.randn(2 * _SPLIT_ROWS + 9, 3, dtype=torch.float64, requires_grad=True)
    Ws = torch.randn(3, 2, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradgradcheck(lambda a, b: _TallMM.apply(a, b), (xs, Ws))
