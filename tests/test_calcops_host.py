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
