"""Writes tests/golden/oracle_small.npz: small input/output vectors of the
oracle (FP64), frozen so that a later change to the oracle or to the library
shows up as a difference against fixed data rather than against a re-run.

Contents (all arrays C-ordered in the reference layout):
  ran1_seed1, ran1_seedm3   first 64 iy values of ran1 (initialize.f90:516-548)
  ic16                      initialize(16^3, iseed=-3, akp=4, u0=1): 3 fields
  rhs16_forced              N(ic16) of dns-deform.f90:259-288, nu=0.01, forcing k^2<4, eps=0.1
  step16_forced             one IF-RK4 step of ic16, dt=0.002
  les16                     LES RHS of ic16 (static c_s^2 = 0.0289, Delta = 2 pi/16)
  ic2d32                    2D initial vorticity 32^2, seed -5, k0=8, energy 0.5
  rhs2d32                   2D RHS of ic2d32 (1e-5 k^8, damping 0.1, band 4..5, eps 0.01)
Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import les_scalar as les  # noqa: E402
from oracle import spectral3d as orc  # noqa: E402
from oracle import vort2d as v2  # noqa: E402


def build():
    n = 16
    out = {"ran1_seed1": orc.ran1_stream(1, 64), "ran1_seedm3": orc.ran1_stream(-3, 64)}
    ic = orc.initialize(n, n, n, -3, 4.0, 1.0)
    out["ic16"] = np.stack(ic)
    d = orc.DNS(orc.DNSParams(nx=n, nu=0.01, dt=0.002, forced=True, fkmax=2.0, eps=0.1))
    out["rhs16_forced"] = np.stack(d.rhs(*ic))
    out["step16_forced"] = np.stack(d.rk4_step(*ic))
    o = les.LESScalar(les.LESParams(nx=n, nu=0.01, dt=0.002, forced=False, cs2=0.0289, delta=2 * np.pi / n,
                                    scalar=False))
    out["les16"] = np.stack(o.rhs(*ic, np.zeros_like(ic[0]))[:3])
    w = v2.initialize2d(32, -5, 8.0, 0.5)
    out["ic2d32"] = w
    out["rhs2d32"] = v2.Vort2D(v2.Vort2DParams(n=32)).rhs(w)
    return out


if __name__ == "__main__":
    path = os.path.join(ROOT, "tests", "golden", "oracle_small.npz")
    np.savez_compressed(path, **build())
    print("wrote", path, os.path.getsize(path), "bytes")
