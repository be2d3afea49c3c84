"""matpl_amd.tallmm: the split-K product and the two-stage bias-gradient sum used by the training
paths equal plain autograd in value, gradient and second derivative (sizes below, at and above the
split thresholds, with padding)."""
import torch

from matpl_amd.tallmm import _SPLIT_ROWS, _SUM_ROWS, _TallMM, add_bias, colsum


def test_tall_products_match_plain_autograd():
    torch.manual_seed(0)
    for K in (100, 2 * _SPLIT_ROWS + 1, 3 * _SPLIT_ROWS + 7):
        x = torch.randn(K, 7, dtype=torch.float64, requires_grad=True)
        W = torch.randn(7, 5, dtype=torch.float64, requires_grad=True)
        outs = []
        for f in (lambda a, b: _TallMM.apply(a, b, False), lambda a, b: a @ b):
            xx, WW = x.detach().clone().requires_grad_(), W.detach().clone().requires_grad_()
            y = torch.tanh(f(xx, WW))
            (gx,) = torch.autograd.grad(y.pow(2).sum(), xx, create_graph=True)
            gW = torch.autograd.grad((gx ** 2).sum(), WW)[0]
            outs.append((y.detach(), gx.detach(), gW))
        for a, b in zip(*outs):
            assert torch.allclose(a, b, rtol=1e-12, atol=1e-10)
    xs = torch.randn(2 * _SPLIT_ROWS + 9, 3, dtype=torch.float64, requires_grad=True)
    Ws = torch.randn(3, 2, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradgradcheck(lambda a, b: _TallMM.apply(a, b, False), (xs, Ws))


def test_bias_add_and_column_sum():
    torch.manual_seed(1)
    for shape in ((5000, 25), (7, 300, 25), (10, 3)):
        g = torch.randn(*shape, dtype=torch.float64)
        assert torch.allclose(colsum(g), g.reshape(-1, shape[-1]).sum(0), rtol=1e-13, atol=1e-12)
    for K in (10, 4 * _SUM_ROWS + 76):
        y = torch.randn(K, 3, dtype=torch.float64, requires_grad=True)
        b = torch.randn(3, dtype=torch.float64, requires_grad=True)
        f = lambda y_, b_: torch.tanh(add_bias(y_, b_))
        assert torch.autograd.gradcheck(f, (y, b))
        assert torch.autograd.gradgradcheck(f, (y, b))
