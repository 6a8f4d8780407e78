"""Argument checks of the torch.autograd wrapper over the adjoint, on the CPU
(no library call is made: every case is rejected before one)."""
import pytest

torch = pytest.importorskip("torch")

import speccodelib_amd as scl  # noqa: E402
from speccodelib_amd import torch_adjoint as ta  # noqa: E402


class _Plan:
    def __init__(self, prec, shape):
        self.prec = prec
        self.shape = shape


class _Adj:
    """Stands in for an Adjoint: any library call through it fails the test."""

    def __init__(self, prec, shape=(16, 16, 9)):
        self.plan = _Plan(prec, shape)
        self._fwd_token = None

    def forward_device(self, *a):
        raise AssertionError("library called with a rejected argument")

    backward_device = forward_device


@pytest.mark.parametrize("prec,dtype", [(scl.DP, torch.complex64), (scl.SP, torch.complex128),
                                        (scl.DP, torch.float64)])
def test_dtype_mismatch_rejected(prec, dtype):
    u0 = torch.zeros((3, 16, 16, 9), dtype=dtype)
    with pytest.raises(ValueError, match="dtype"):
        ta.rk4_map(u0, 0.0225, _Adj(prec))


def test_shape_mismatch_rejected():
    u0 = torch.zeros((3, 16, 16, 8), dtype=torch.complex128)
    with pytest.raises(ValueError, match="shape"):
        ta.rk4_map(u0, 0.0225, _Adj(scl.DP))
    u0 = torch.zeros((2, 16, 16, 9), dtype=torch.complex128)
    with pytest.raises(ValueError, match="shape"):
        ta.rk4_map(u0, 0.0225, _Adj(scl.DP))


def test_host_tensor_rejected():
    u0 = torch.zeros((3, 16, 16, 9), dtype=torch.complex128)
    with pytest.raises(ValueError, match="CUDA"):
        ta.rk4_map(u0, 0.0225, _Adj(scl.DP))
