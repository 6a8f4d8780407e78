"""GPU parity of the long z-line kernels against the oracle: grids whose
z-lines have 512 and 1024 points (the 512^3 and 1024^3 runs' z-transforms,
in 1024-thread workgroups that keep one transformed component in LDS) on
small x-y planes, so the FP64 oracle finishes in seconds; FP64 plans (four-
column tiles at 1024 points) on 512- and 1024-point z- and y-lines.

Inputs are white-noise spectra over every mode (Hermitian on kx = 0), so
every z-wavenumber is exercised.  ratesdudt (dns-deform.f90:259-288) runs
the z-inverse and the mode-0 z-forward; two IF-RK4 steps
(dns-deform.f90:160-237) run the four stage variants.  Bar: 2e-5 of the
largest entry (FP32 rounding times log2 1024) against the FP64 oracle; 1e-11
for FP64 plans."""
import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu

CASES = [((32, 32, 512), scl.SP), ((32, 16, 1024), scl.SP), ((32, 32, 512), scl.DP), ((16, 512, 16), scl.DP),
         ((32, 16, 1024), scl.DP), ((16, 1024, 16), scl.DP)]


def relerr(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def white(shape, seed, k2):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(3):
        a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / (1.0 + k2)
        orc.symmetrize3d(a)
        out.append(a)
    return out


@pytest.mark.parametrize("dims,prec", CASES)
def test_long_z_lines_rhs_and_step(dims, prec):
    bar = 2e-5 if prec == scl.SP else 1e-11
    nx, ny, nz = dims
    print(f"case long z-lines {dims} prec={prec}", flush=True)
    p = scl.Plan3D(nx, ny, nz, prec)
    op = orc.DNSParams(nx=nx, ny=ny, nz=nz, nu=0.01, dt=0.002, forced=True, fkmax=2.0, eps=0.1)
    d = orc.DNS(op)
    v64 = white(p.shape, 11, d.k2.astype(np.float64))
    cd = scl._cdtype(prec)
    v = [a.astype(cd) for a in v64]
    # the oracle starts from the same rounded input
    v64 = [a.astype(np.complex128) for a in v]
    want = d.rhs(*v64)
    got = p.ratesdudt(*v, forced=1, fkmax=2.0, eps=0.1, eta=0.0)
    err = max(relerr(a, b) for a, b in zip(got, want))
    print(f"ratesdudt rel err {err:.2e}", flush=True)
    assert err < bar
    w = tuple(v64)
    u = [a.copy() for a in v]
    for _ in range(2):
        w = d.rk4_step(*w)
        p.rk4step(*u, 0.01, 0.002, 1, 2.0, 0.1)
    err = max(relerr(a, b) for a, b in zip(u, w))
    print(f"2 IF-RK4 steps rel err {err:.2e}", flush=True)
    assert err < bar
    p.close()
