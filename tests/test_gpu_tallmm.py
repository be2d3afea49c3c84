"""GPU: the tall-by-small product kernel (matpl_tall_mm) against a float64 product of the same operands.
Bounds: an FP32 sum of a products is within a * 2^-23 * sum|x_j W_jc| of the exact value (FP32 mode);
in lowp mode the operands are bfloat16 values, the result is a bfloat16 value within one bfloat16 ulp
plus that FP32 sum bound of the exact product, and within two ulps plus twice the bound of torch's BF16
product with FP32 accumulation. Ragged row counts, a, b from 1 to 64; autograd through
tall_mm_f32 / tall_mm_bf16 (first and second order) matches the float64 product's."""
import pytest
import torch

from matpl_amd.tallmm import _small_mm, tall_mm_bf16, tall_mm_f32

pytestmark = pytest.mark.gpu

SHAPES = [(0, 25, 25), (1, 1, 25), (63, 25, 25), (64, 25, 1), (1000, 25, 50), (4097, 50, 50), (777, 64, 64),
          (130, 1, 50), (65, 50, 1), (20000, 16, 3)]


@pytest.mark.parametrize("K,a,b", SHAPES)
def test_fp32_matches_float64(dev, K, a, b):
    print(f"case tall_mm fp32 K={K} a={a} b={b}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(K * 131 + a * 7 + b)
    x = torch.randn(K, a, generator=g).to(dev)
    W = torch.randn(a, b, generator=g).to(dev)
    y = _small_mm(x, W, False)
    torch.cuda.synchronize()
    ref = x.double() @ W.double()
    assert y.shape == (K, b) and y.dtype == torch.float32
    if K:
        scale = (x.double().abs() @ W.double().abs()).clamp_min(1e-30)
        assert float(((y.double() - ref).abs() / scale).max()) <= a * 2.0 ** -23


@pytest.mark.parametrize("K,a,b", SHAPES)
def test_bf16_matches_torch_bf16(dev, K, a, b):
    print(f"case tall_mm bf16 K={K} a={a} b={b}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(K * 17 + a * 3 + b)
    x = torch.randn(K, a, generator=g).to(dev)
    W = torch.randn(a, b, generator=g).to(dev)
    y = _small_mm(x, W, True)
    torch.cuda.synchronize()
    if not K:
        return
    xb, Wb = x.to(torch.bfloat16).double(), W.to(torch.bfloat16).double()
    exact = xb @ Wb
    sumerr = a * 2.0 ** -23 * (xb.abs() @ Wb.abs())
    ulp = torch.exp2(torch.floor(torch.log2(exact.abs().clamp_min(1e-30))) - 7)
    assert torch.equal(y, y.to(torch.bfloat16).float())                      # the result is a bf16 value
    assert bool(((y.double() - exact).abs() <= ulp + sumerr).all())
    tb = (x.to(torch.bfloat16) @ W.to(torch.bfloat16)).float()
    assert bool(((y.double() - tb.double()).abs() <= 2 * ulp + 2 * sumerr).all())


@pytest.mark.parametrize("fn,tol", [(tall_mm_f32, 1e-5), (tall_mm_bf16, 3e-2)])
def test_autograd_matches_torch(dev, fn, tol):
    print(f"case tall_mm autograd {fn.__name__}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(5)
    x = torch.randn(3000, 25, generator=g).to(dev).requires_grad_()
    W = torch.randn(25, 50, generator=g).to(dev).requires_grad_()
    up = torch.randn(3000, 50, generator=g).to(dev)
    y = fn(x, W)
    gx, gW = torch.autograd.grad((y * up).sum(), (x, W), create_graph=True)
    ggW = torch.autograd.grad((gx ** 2).sum(), W)[0]
    x2, W2 = x.detach().double().requires_grad_(), W.detach().double().requires_grad_()
    y2 = x2 @ W2
    gx2, gW2 = torch.autograd.grad((y2 * up.double()).sum(), (x2, W2), create_graph=True)
    ggW2 = torch.autograd.grad((gx2 ** 2).sum(), W2)[0]
    for a, b in ((y, y2), (gx, gx2), (gW, gW2), (ggW, ggW2)):
        assert float((a.detach().double() - b.detach()).norm() / b.detach().norm()) <= tol


@pytest.mark.parametrize("lowp", [False, True])
def test_strided_operands(dev, lowp):
    """A transposed weight view and an x that is a column slice of a wider tensor (row stride > a)
    give exactly the result of the same product on contiguous copies."""
    print(f"case tall_mm strided lowp={lowp}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(11)
    wide = torch.randn(2001, 70, generator=g).to(dev)
    x = wide[:, 3:28]                                   # [2001, 25], row stride 70
    Wt = torch.randn(50, 25, generator=g).to(dev)
    W = Wt.t()                                          # [25, 50] view, strides (1, 25)
    y = _small_mm(x, W, lowp)
    y0 = _small_mm(x.contiguous(), W.contiguous(), lowp)
    torch.cuda.synchronize()
    assert torch.equal(y, y0)
