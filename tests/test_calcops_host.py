"""Host-side argument checks of the CalcOps_cuda operators (no GPU needed): tensors off the GPU and
unsupported dtypes raise before anything is launched, as the reference's operators reject them
(src/op/src/CalcOps.cpp checks device and dtype before launching)."""
import pytest
import torch

from matpl_amd import calcops


def test_cpu_tensors_are_rejected():
    nl = torch.zeros((1, 4, 6), dtype=torch.int32)
    dE = torch.zeros((1, 4, 24), dtype=torch.float64)
    Rid = torch.zeros((1, 4, 24, 3), dtype=torch.float64)
    F = torch.zeros((1, 4, 3), dtype=torch.float64)
    with pytest.raises(RuntimeError, match="on the GPU"):
        calcops.dp_force(nl, dE, Rid, F, 0)
    with pytest.raises(RuntimeError, match="on the GPU"):
        calcops.dp_virial(nl, dE, torch.zeros((1, 4, 6, 3), dtype=torch.float64), Rid, 0)


@pytest.mark.parametrize("dtype", [torch.int32, torch.float16, torch.bfloat16, torch.int64])
def test_unsupported_dtypes_are_rejected(dtype):
    with pytest.raises(RuntimeError, match="unsupported dtype"):
        calcops._sfx(torch.zeros(1, dtype=dtype))


def test_nghost_from_int_or_cpu_tensor():
    """nghost as an int or a CPU scalar tensor (the pair style's form) is read without a device sync."""
    assert calcops._nghost(7) == 7
    assert calcops._nghost(torch.tensor(7)) == 7


def test_nep_op_checks_cached_per_tensor_version():
    """The NEP ops' host-side input checks (largest index, atom count, frame index) are read once per
    tensor and version: an in-place change of the tensor invalidates them, and inference-mode
    tensors (no version counter) are checked every time."""
    from matpl_amd import nep_ops
    t = torch.tensor([3, 5, 2])
    assert nep_ops._int_max(t) == 5 and nep_ops._int_sum(t) == 10
    calls = []
    assert nep_ops._cached(t, "k", lambda: calls.append(1) or 7) == 7
    assert nep_ops._cached(t, "k", lambda: calls.append(1) or 8) == 7 and len(calls) == 1
    t[0] = 9
    assert nep_ops._int_max(t) == 9 and nep_ops._int_sum(t) == 16
    assert nep_ops._cached(t, "k", lambda: calls.append(1) or 8) == 8 and len(calls) == 2
    start, af = nep_ops._frame_index(torch.tensor([2, 3]), torch.device("cpu"))
    assert start.tolist() == [0, 2, 5] and af.tolist() == [0, 0, 1, 1, 1]
    with torch.inference_mode():
        u = torch.tensor([1, 4])
        assert nep_ops._int_max(u) == 4
        u[1] = 0
        assert nep_ops._int_max(u) == 1
