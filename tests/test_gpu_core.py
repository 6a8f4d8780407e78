"""GPU parity of the building blocks (through the C ABI) against the oracle:
3D transforms (fftwplan.f90 conventions), convec_dns, symmetrize,
dealiasing, projection, force, force_heli and the full ratesdudt RHS.

Tolerances: FP32 results agree with the FP64 oracle to 3e-6 of the field's
max magnitude (a few units of FP32 rounding times log2 N); FP64 plans agree
to 1e-12.  Index/zeroing work (symmetrize, dealiasing) is bit-exact."""
import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu

SIZES = [(16, 16, 16), (32, 32, 32), (64, 64, 64), (32, 16, 64), (128, 128, 128)]


def tol(prec):
    return 3e-6 if prec == scl.SP else 1e-12


def relerr(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rand_field(shape, seed, prec):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return a.astype(scl._cdtype(prec))


@pytest.mark.parametrize("prec", [scl.SP, scl.DP])
@pytest.mark.parametrize("dims", SIZES)
def test_r2c_c2r_match_oracle(dims, prec):
    nx, ny, nz = dims
    print(f"case r2c/c2r {dims} prec={prec}", flush=True)
    p = scl.Plan3D(nx, ny, nz, prec)
    rng = np.random.default_rng(1)
    r = rng.standard_normal((nz, ny, nx))
    a = p.packed_real(r.astype(scl._rdtype(prec)))
    p.r2c(a)
    ref = orc.r2c(r)
    assert relerr(a, ref) < tol(prec)
    # c2r of a Hermitian-consistent spectrum
    b = ref.astype(scl._cdtype(prec))
    out = p.c2r(b)
    assert np.abs(out - r * nx * ny * nz).max() / (nx * ny * nz * np.abs(r).max()) < tol(prec)
    # c2r of an arbitrary half spectrum (Im of kx=0 and kx=nx/2 ignored)
    c = rand_field(p.shape, 2, prec)
    want = orc.c2r(c.astype(np.complex128), nx)
    got = p.c2r(c.copy())
    assert np.abs(got - want).max() / np.abs(want).max() < tol(prec)


@pytest.mark.parametrize("prec", [scl.SP, scl.DP])
def test_transforms_are_deterministic(prec):
    p = scl.Plan3D(64, 64, 64, prec)
    c = rand_field(p.shape, 3, prec)
    a = p.c2r(c.copy()).copy()
    for _ in range(3):
        np.testing.assert_array_equal(p.c2r(c.copy()), a)


@pytest.mark.parametrize("prec", [scl.SP, scl.DP])
@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 16, 64), (64, 64, 64)])
def test_convec_dns(dims, prec):
    nx, ny, nz = dims
    print(f"case convec_dns {dims} prec={prec}", flush=True)
    p = scl.Plan3D(nx, ny, nz, prec)
    f = [rand_field(p.shape, s, prec) for s in range(6)]
    for a in f:  # realisable fields: Hermitian on kx = 0
        orc.symmetrize3d(a)
    want = orc.convec_dns(*[a.astype(np.complex128) for a in f], nx, ny, nz)
    v = [a.copy() for a in f[:3]]
    w = [a.copy() for a in f[3:]]
    p.convec_dns(*v, *w)
    for a, b in zip(w, want):
        assert relerr(a, b) < 3 * tol(prec)
    for a, b in zip(v, f[:3]):  # intent(in)
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("prec", [scl.SP, scl.DP])
@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 16, 64)])
def test_symmetrize_and_dealiasing_bit_exact(dims, prec):
    nx, ny, nz = dims
    p = scl.Plan3D(nx, ny, nz, prec)
    c = rand_field(p.shape, 5, prec)
    want = orc.symmetrize3d(c.copy())
    got = c.copy()
    p.symmetrize(got)
    np.testing.assert_array_equal(got, want)
    kx, ky, kz, k2 = orc.wavenumber(nx, ny, nz, "sp" if prec == scl.SP else "dp")
    want = orc.dealiasing(c.copy(), k2)
    got = c.copy()
    p.dealiasing(got, k2)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("prec", [scl.SP, scl.DP])
def test_projection(prec):
    p = scl.Plan3D(32, 32, 32, prec)
    pr = "sp" if prec == scl.SP else "dp"
    kx, ky, kz, k2 = orc.wavenumber(32, 32, 32, pr)
    v = [rand_field(p.shape, s, prec) for s in range(3)]
    want = [a.copy() for a in v]
    orc.projection(*want, kx, ky, kz, pr)
    p.projection(*v, kx, ky, kz)
    for a, b in zip(v, want):
        assert relerr(a, b) < tol(prec)


def test_wavenumber_matches_oracle():
    kx, ky, kz, k2 = scl.wavenumber(32, 16, 64, scl.SP)
    okx, oky, okz, ok2 = orc.wavenumber(32, 16, 64, "sp")
    for a, b in [(kx, okx), (ky, oky), (kz, okz), (k2, ok2)]:
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("prec", [scl.SP, scl.DP])
def test_force_and_force_heli(prec):
    n = 32
    pr = "sp" if prec == scl.SP else "dp"
    p = scl.Plan3D(n, n, n, prec)
    kx, ky, kz, k2 = orc.wavenumber(n, n, n, pr)
    v = list(orc.initialize(n, n, n, -3, 4.0, 1.0, prec=pr))
    w = [rand_field(p.shape, s, prec) for s in range(3)]
    want = [a.astype(np.complex128) for a in w]
    orc.force(*want, *[a.astype(np.complex128) for a in v], orc.wavenumber(n, n, n)[3], 2.0, 0.1)
    got = [a.copy() for a in w]
    p.force(*got, *v, k2, 2.0, 0.1)
    for a, b in zip(got, want):
        assert relerr(a, b) < tol(prec)
    om = [a.astype(scl._cdtype(prec)) for a in orc.vorticity(*v, kx, ky, kz)]
    want = orc.force_heli(*[a.astype(np.complex128) for a in v], *[a.astype(np.complex128) for a in om],
                          orc.wavenumber(n, n, n)[3], 2.5, 0.1, 0.02)
    got = p.force_heli(*v, *om, k2, 2.5, 0.1, 0.02)
    for a, b in zip(got, want):
        assert relerr(a, b) < 10 * tol(prec)


@pytest.mark.parametrize("prec", [scl.SP, scl.DP])
@pytest.mark.parametrize("forced", [0, 1, 2])
def test_ratesdudt(prec, forced):
    n = 32
    pr = "sp" if prec == scl.SP else "dp"
    print(f"case ratesdudt n={n} prec={prec} forced={forced}", flush=True)
    p = scl.Plan3D(n, n, n, prec)
    v = orc.initialize(n, n, n, 1, 4.0, 1.0, prec=pr)
    d = orc.DNS(orc.DNSParams(nx=n, forced=forced == 1, fkmax=2.0, eps=0.1))
    v64 = [a.astype(np.complex128) for a in v]
    if forced == 2:
        # the helical RHS: w = u x w + force_heli(u, w), then symmetrize + projection
        om = orc.vorticity(*v64, d.kx, d.ky, d.kz)
        nl = orc.convec_dns(*v64, *om, n, n, n)
        f = orc.force_heli(*v64, *om, d.k2, 2.0, 0.1, 0.05)
        want = [a + b for a, b in zip(nl, f)]
        for a in want:
            orc.symmetrize3d(a)
        orc.projection(*want, d.kx, d.ky, d.kz)
    else:
        want = d.rhs(*v64)
    got = p.ratesdudt(*v, forced=forced, fkmax=2.0, eps=0.1, eta=0.05)
    for a, b in zip(got, want):
        assert relerr(a, b) < 3 * tol(prec)
