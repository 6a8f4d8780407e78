"""The LES / scalar and adjoint parity tests again with the separate prep and
combine passes forced at small sizes (SCL_EXT_UNFUSED=1): the path plans of
1024-point z-lines take, where the fused z-passes of those RHS would spill.
It runs the prep kernels, the z-passes of the plain column transform, the
x-pass with the physical products, the separate combine kernels (IF-RK4
and IF-AB2), and the
adjoint's prep, pointwise and r2c passes without the physical tape, against
the same oracle bars."""
import pytest

import speccodelib_amd as scl
from tests import test_gpu_ab2_les as ab2les
from tests import test_gpu_adjoint as adj
from tests import test_gpu_les as les

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def unfused(monkeypatch):
    monkeypatch.setenv("SCL_EXT_UNFUSED", "1")


@pytest.mark.parametrize("prec", [scl.SP, scl.DP])
def test_convecphi(prec):
    les.test_convecphi_dns(prec)


def test_les_rhs_and_solver_fp64():
    les.test_les_scalar_rhs_fp64()
    for dynamic in (False, True):
        les.test_les_scalar_solver_fp64(dynamic)


def test_les_scalar_ab2_fp64():
    for dynamic in (False, True):
        ab2les.test_les_scalar_ab2_fp64(dynamic)


@pytest.mark.parametrize("sgs", [0, 1])
def test_adjoint_fp64(sgs):
    adj.test_vjp_dot_product_fp64(sgs)
    adj.test_gradient_vs_fd_fp64_small(sgs)


def test_physical_tape_needs_fused_z_passes(monkeypatch):
    """SCL_ADJ_PHYS_TAPE=1 is refused on the unfused path (the products that
    read the tape use the fused z-inverse)."""
    monkeypatch.setenv("SCL_ADJ_PHYS_TAPE", "1")
    p = scl.Plan3D(32, 32, 32, scl.DP)
    with pytest.raises(Exception, match="physical tape"):
        scl.Adjoint(p, 2, 2, nu=0.005, dt=0.002, sgs=1)
    p.close()
