"""FP64 GPU results against the frozen vectors of tests/golden/oracle_small.npz
(tests/golden/make_golden.py): RHS, one IF-RK4 step, LES RHS (through the
derivs form, N - nu k^2 v), the 2D initial field and 2D RHS; bar 1e-11 of
the field's max."""
import os

import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_small.npz")


def close(a, b, tol=1e-11):
    a = np.asarray(a, np.complex128)
    return np.abs(a - b).max() <= tol * np.abs(b).max()


def test_3d_against_golden():
    g = np.load(GOLD)
    n = 16
    print("case golden 3d fp64 16^3", flush=True)
    p = scl.Plan3D(n, n, n, scl.DP)
    ic = [np.ascontiguousarray(x) for x in g["ic16"]]
    assert close(np.stack(p.ratesdudt(*[x.copy() for x in ic], forced=1, fkmax=2.0, eps=0.1)), g["rhs16_forced"])
    v = [x.copy() for x in ic]
    p.rk4step(*v, 0.01, 0.002, 1, 2.0, 0.1)
    assert close(np.stack(v), g["step16_forced"])
    d = scl.DNS(p, nu=0.01, dt=0.002, forced=0, sgs=1, cs2=0.0289, delta=2 * np.pi / n)
    k2 = orc.wavenumber(n, n, n)[3].astype(np.float64)
    k2[0, 0, 0] = 0
    got = np.stack(d.rhs(*[x.copy() for x in ic]))
    assert close(got, g["les16"] - 0.01 * k2 * g["ic16"])
    d.close()
    p.close()


def test_2d_against_golden():
    g = np.load(GOLD)
    print("case golden 2d fp64 32^2", flush=True)
    e = scl.Ensemble2D([(32, 7)], prec=scl.DP)
    e.initialize(seed0=-1, k0=8.0, energy=0.5)  # member 4 draws from seed -5
    assert close(e.get_member(0, 4), g["ic2d32"], 1e-13)
    assert close(e.rhs(0, 2, np.ascontiguousarray(g["ic2d32"])), g["rhs2d32"])
    e.close()
