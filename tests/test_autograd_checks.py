"""Argument checks of the differentiable solve (spand_public_amd.autograd): every tensor that
the library reads as device memory is checked for device, dtype and shape first, so that a
wrong argument is a ValueError and never a read of host memory or past a buffer's end."""
import pytest

torch = pytest.importorskip("torch")

from spand_public_amd import autograd as ad  # noqa: E402


def test_rhs_checks():
    N = 10
    with pytest.raises(ValueError, match="CUDA"):
        ad.check_rhs(torch.zeros(N, 2, dtype=torch.float64), N)  # host tensor
    with pytest.raises(ValueError, match="CUDA"):
        ad.check_rhs([0.0] * N, N)
    meta = torch.empty(N, 2, dtype=torch.float64, device="meta")
    with pytest.raises(ValueError):
        ad.check_rhs(meta, N)  # not a CUDA tensor either


def test_value_checks():
    with pytest.raises(ValueError, match="CUDA"):
        ad.check_values(torch.zeros(7, dtype=torch.float64), 7)
    with pytest.raises(ValueError):
        ad.check_values(None, 7)


def test_solver_kind_is_checked():
    with pytest.raises(ValueError, match="symm"):
        ad.SparseSolver([0, 1], [0], [[0.0]], 1, symm="SYM")
