"""Diagnostics on the GPU: the skewness (skewness.f90:1-48) against the
oracle, host-array and solver forms; E(k), H(k) and the totals bitwise the
same on repeated calls (fixed summation order), also from a second solver
holding the same state; spectra on non-cubic grids and at the 1024-point
row length (whose shells span 9 passes of a wavefront per row) against the
oracle."""
import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,n", [(scl.DP, 32), (scl.SP, 64)])
def test_skewness_matches_oracle(prec, n):
    print(f"case skewness prec={prec} n={n}", flush=True)
    p = scl.Plan3D(n, n, n, prec)
    v = orc.initialize(n, n, n, -3, 4.0, 1.0)
    kx = orc.wavenumber(n, n, n)[0]
    # a skewed field: the seeded one plus a quadratic term of its own
    u = orc.c2r(v[0], n)
    v0 = v[0] + 0.3 * orc.r2c(u * u) / n**3
    orc.symmetrize3d(v0)
    want = orc.skewness(v0, kx, n)
    cd = np.complex128 if prec == scl.DP else np.complex64
    rd = np.float64 if prec == scl.DP else np.float32
    got = p.skewness(v0.astype(cd), kx.astype(rd))
    tol = 1e-12 if prec == scl.DP else 2e-5
    print("skewness", got, want, flush=True)
    assert abs(got - want) <= tol * max(1.0, abs(want))
    s = scl.DNS(p)
    s.set_velocity(v0.astype(cd), v[1].astype(cd), v[2].astype(cd))
    assert abs(s.diagnostics()["skewness"] - got) <= 1e-13 * max(1.0, abs(got))
    s.close()
    p.close()


def test_diagnostics_bitwise_repeatable():
    n = 64
    print("case diagnostics repeat", flush=True)
    p = scl.Plan3D(n, n, n, scl.SP)
    s = scl.DNS(p, nu=0.01, dt=0.002, forced=1, fkmax=2.0, eps=0.1)
    s.initialize(1, 4.0, 1.0)
    s.step(3)
    s.sync()
    ek, hk, dg = s.spectrum(), s.helicity_spectrum(), s.diagnostics()
    v = s.get_velocity()
    s2 = scl.DNS(p)
    s2.set_velocity(*v)
    for _ in range(3):
        for d in (s, s2):
            assert np.array_equal(d.spectrum(), ek) and np.array_equal(d.helicity_spectrum(), hk)
            dg2 = d.diagnostics()
            for k in ("energy", "dissipation", "enstrophy", "helicity", "helicity_dissipation", "skewness",
                      "divrms_spec"):
                assert dg2[k] == dg[k], k
    s.close()
    s2.close()
    p.close()


@pytest.mark.parametrize("shape", [(1024, 16, 16), (16, 64, 32), (256, 32, 16)])
def test_spectra_noncubic_and_long_rows(shape):
    nx, ny, nz = shape
    print(f"case spectra shape={shape}", flush=True)
    p = scl.Plan3D(nx, ny, nz, scl.DP if nx <= 256 else scl.SP)
    rng = np.random.default_rng(3)
    sh = (nz, ny, nx // 2 + 1)
    v = [rng.standard_normal(sh) + 1j * rng.standard_normal(sh) for _ in range(3)]
    kx, ky, kz, k2 = orc.wavenumber(nx, ny, nz)
    cd = np.complex128 if p.prec == scl.DP else np.complex64
    v = [a.astype(cd) for a in v]
    s = scl.DNS(p)
    s.set_velocity(*v)
    e_ref = orc.spectrum(*v, k2)[: nx // 2]
    h_ref = orc.helicity_spectrum(*v, kx, ky, kz, k2)[: nx // 2]
    tol = 1e-12
    np.testing.assert_allclose(s.spectrum(), e_ref, rtol=tol, atol=tol * np.abs(e_ref).max())
    np.testing.assert_allclose(s.helicity_spectrum(), h_ref, rtol=tol, atol=tol * np.abs(h_ref).max())
    s.close()
    p.close()
