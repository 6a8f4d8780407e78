"""The GPU parity tests again with the z-blocked row order forced at small
sizes: rows ordered [z/8][y][z mod 8] (SCL_ZBLOCK=8), the order the plans
of 1024-point z-lines use, against the same oracle bars and fixtures.
Every kernel that addresses rows by (y, z) runs here: the column tiles of
the y- and z-passes, the fused z-pass (both template variants of the row
order), the Hermitian copy and its adjoint, the forcing box, the entry
decode of the pointwise, LES, adjoint and diagnostic kernels, and the
reference-layout conversions at the boundary."""
import pytest

import speccodelib_amd as scl
from tests import test_gpu_adjoint as adj
from tests import test_gpu_core as core
from tests import test_gpu_derivs as der
from tests import test_gpu_golden as gold
from tests import test_gpu_les as les
from tests import test_gpu_rk4 as rk4
from tests import test_gpu_zlines as zl

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def zblock8(monkeypatch):
    monkeypatch.setenv("SCL_ZBLOCK", "8")


@pytest.mark.parametrize("prec", [scl.SP, scl.DP])
def test_transforms(prec):
    core.test_r2c_c2r_match_oracle((32, 16, 64), prec)
    core.test_r2c_c2r_match_oracle((16, 16, 16), prec)


@pytest.mark.parametrize("prec", [scl.SP, scl.DP])
def test_convec_symmetrize_dealiasing(prec):
    core.test_convec_dns((32, 16, 64), prec)
    core.test_symmetrize_and_dealiasing_bit_exact((32, 16, 64), prec)


@pytest.mark.parametrize("prec", [scl.SP, scl.DP])
def test_force_and_rhs(prec):
    core.test_force_and_force_heli(prec)
    for forced in (0, 1, 2):
        core.test_ratesdudt(prec, forced)


def test_rk4_solver_fp64():
    for forced in (0, 1):
        rk4.test_rk4step_fp64_matches_oracle(forced)
    for g in (False, True):
        rk4.test_device_solver_fp64(g)
    rk4.test_helical_forcing_solver_fp64()
    rk4.test_diagnostics_match_oracle_fp64()


def test_golden_3d():
    gold.test_3d_against_golden()


def test_derivs():
    der.test_dns_rhs_matches_oracle()
    der.test_ifrk4_driver_with_builtin_callback()


def test_les_fp64():
    les.test_les_scalar_rhs_fp64()
    les.test_smag_and_static_dynsmag_fp64()
    les.test_dynsmag_update_fp64()
    for dynamic in (False, True):
        les.test_les_scalar_solver_fp64(dynamic)


def test_adjoint_fp64():
    for sgs in (0, 1):
        adj.test_vjp_dot_product_fp64(sgs)
        adj.test_gradient_vs_fd_fp64_small(sgs)


def test_long_z_lines_512():
    zl.test_long_z_lines_rhs_and_step((32, 32, 512), scl.SP)
    zl.test_long_z_lines_rhs_and_step((32, 32, 512), scl.DP)
