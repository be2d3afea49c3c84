"""Batched small products (csrc/bsmm.hip via matpl_amd.bsmm): values against torch.matmul, first and
second derivatives by gradcheck / gradgradcheck in float64, float32 at float32 precision, and the
size limit."""
import pytest
import torch

from matpl_amd import _lib
from matpl_amd.bsmm import bmm_nn, bmm_nt, bmm_tn

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,K,a,b", [(7, 36, 4, 25), (3, 4, 25, 16), (5, 1, 3, 2), (130, 9, 1, 7)])
def test_values(dev, n, K, a, b):
    g = torch.Generator(device="cpu").manual_seed(n + K)
    X = torch.randn(n, K, a, generator=g, dtype=torch.float64).to(dev)
    Y = torch.randn(n, K, b, generator=g, dtype=torch.float64).to(dev)
    Z = torch.randn(n, a, b, generator=g, dtype=torch.float64).to(dev)
    ref = {"tn": X.transpose(1, 2) @ Y, "nt": Y @ Z.transpose(1, 2).contiguous(), "nn": X @ Z}
    got = {"tn": bmm_tn(X, Y), "nt": bmm_nt(Y, Z), "nn": bmm_nn(X, Z)}
    for k in ref:
        assert torch.allclose(got[k], ref[k], rtol=1e-13, atol=1e-12), k
    got32 = {"tn": bmm_tn(X.float(), Y.float()), "nt": bmm_nt(Y.float(), Z.float()), "nn": bmm_nn(X.float(), Z.float())}
    for k in ref:
        assert got32[k].dtype == torch.float32
        assert torch.allclose(got32[k].double(), ref[k], rtol=1e-5, atol=1e-5 * K), k


def test_derivatives(dev):
    g = torch.Generator(device="cpu").manual_seed(3)
    mk = lambda *s: torch.randn(*s, generator=g, dtype=torch.float64).to(dev).requires_grad_()
    X, Y, Z = mk(3, 5, 2), mk(3, 5, 4), mk(3, 2, 4)
    for f, args in ((bmm_tn, (X, Y)), (bmm_nt, (Y, Z)), (bmm_nn, (X, Z))):
        h = lambda p, q, f=f: torch.tanh(f(p, q))
        assert torch.autograd.gradcheck(h, args)
        assert torch.autograd.gradgradcheck(h, args)


def test_size_limit(dev):
    X = torch.zeros(2, 100, 64, dtype=torch.float64, device=dev)
    Y = torch.zeros(2, 100, 64, dtype=torch.float64, device=dev)
    with pytest.raises(_lib.MatplError):
        bmm_tn(X, Y)
