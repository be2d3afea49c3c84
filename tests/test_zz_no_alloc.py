"""matpl_dp_virial_noalloc (calculateVirial with its reduction scratch in a static device array, no
workspace argument, no allocation) against matpl_dp_virial_f32 / _f64 with a caller workspace, on two
calls in a row, the second with a smaller input (fewer frames, atoms and slots), in both precisions;
more frames than the scratch holds is an argument error. The per-atom rows are summed with atomics
(as the reference's calculateVirial.cu does), so two calls of either entry point agree to rounding:
FP64 within 1e-12 and FP32 within 2e-5 of the largest entry."""
import numpy as np
import pytest
import torch

from matpl_amd import _lib
from matpl_amd.calcops import _ptr, _stream

pytestmark = pytest.mark.gpu


def _inputs(dev, dtype, B, N, S, ng, seed):
    rng = np.random.default_rng(seed)
    nl = rng.integers(-1, N + ng, size=(B, N, S)).astype(np.int32)
    T = lambda a: torch.as_tensor(a, dtype=dtype, device=dev).contiguous()
    return (torch.as_tensor(nl, device=dev), T(rng.normal(size=(B, N, S * 4))), T(rng.normal(size=(B, N, S, 3))),
            T(rng.normal(size=(B, N, S * 4, 3))))


def _with_workspace(nl, dE, Rij, Rid, B, N, S, ng):
    v = torch.empty((B, 9), dtype=dE.dtype, device=dE.device)
    av = torch.empty((B, N + ng, 9), dtype=dE.dtype, device=dE.device)
    ws_elems = _lib.query("matpl_dp_virial_workspace_elems", B, N, ng)
    ws = torch.empty(max(ws_elems, 1), dtype=dE.dtype, device=dE.device)
    sfx = "f64" if dE.dtype == torch.float64 else "f32"
    _lib.call(f"matpl_dp_virial_{sfx}", _ptr(nl), _ptr(dE), _ptr(Rij), _ptr(Rid), B, N, S, ng, _ptr(v), _ptr(av),
              _ptr(ws), ws.numel(), _stream())
    return v, av


def _noalloc(nl, dE, Rij, Rid, B, N, S, ng):
    v = torch.full((B, 9), float("nan"), dtype=dE.dtype, device=dE.device)
    av = torch.full((B, N + ng, 9), float("nan"), dtype=dE.dtype, device=dE.device)
    _lib.call("matpl_dp_virial_noalloc", int(dE.dtype == torch.float64), _ptr(nl), _ptr(dE), _ptr(Rij), _ptr(Rid),
              B, N, S, ng, _ptr(v), _ptr(av), _stream())
    return v, av


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_virial_equals_workspace_virial(dev, dtype):
    for k, (B, N, S, ng) in enumerate(((6, 5000, 24, 40), (2, 700, 16, 3))):
        print(f"case virial noalloc {dtype} call {k}: B={B} N={N} S={S} nghost={ng}", flush=True)
        args = _inputs(dev, dtype, B, N, S, ng, seed=11 + k)
        v0, av0 = _with_workspace(*args, B, N, S, ng)
        v1, av1 = _noalloc(*args, B, N, S, ng)
        torch.cuda.synchronize()
        tol = 1e-12 if dtype == torch.float64 else 2e-5
        assert torch.isfinite(v1).all() and torch.isfinite(av1).all(), k
        assert (v1 - v0).abs().max() <= tol * v0.abs().max(), (k, (v1 - v0).abs().max())
        assert (av1 - av0).abs().max() <= tol * av0.abs().max(), (k, (av1 - av0).abs().max())


def test_noalloc_virial_rejects_too_many_frames(dev):
    args = _inputs(dev, torch.float64, 1, 4, 2, 0, seed=1)
    with pytest.raises(_lib.MatplError, match="at most 64"):
        _lib.call("matpl_dp_virial_noalloc", 1, _ptr(args[0]), _ptr(args[1]), _ptr(args[2]), _ptr(args[3]), 65, 4,
                  2, 0, _ptr(args[1]), _ptr(args[1]), _stream())
