"""Pins the oracle's FFT convention (fftwplan.f90:7-15: in-place, unnormalised,
r2c with exp(-ik.x), c2r with exp(+ik.x)) against brute-force DFT sums."""
import numpy as np
import pytest


def brute_r2c(x):
    nz, ny, nx = x.shape
    lx1 = nx // 2 + 1
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = np.zeros((nz, ny, lx1), complex)
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(lx1):
                ph = np.exp(-2j * np.pi * (kx * xx / nx + ky * yy / ny + kz * zz / nz))
                out[kz, ky, kx] = (x * ph).sum()
    return out


def brute_c2r(c, nx):
    """Full Hermitian-extended inverse sum; entries with kx in (0, nx/2) count
    twice via their conjugate partner, kx = 0 and nx/2 count once (real part)."""
    nz, ny, lx1 = c.shape
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = np.zeros((nz, ny, nx))
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(lx1):
                ph = np.exp(2j * np.pi * (kx * xx / nx + ky * yy / ny + kz * zz / nz))
                w = 1.0 if kx in (0, nx // 2) else 2.0
                out += w * np.real(c[kz, ky, kx] * ph)
    return out


@pytest.mark.parametrize("shape", [(4, 6, 8), (8, 8, 8)])
def test_r2c_matches_brute_force(oracle, shape):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(shape)
    np.testing.assert_allclose(oracle.r2c(x), brute_r2c(x), atol=1e-10)


def test_c2r_matches_brute_force_on_hermitian_input(oracle):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((4, 6, 8))
    c = oracle.r2c(x)
    # Hermitian input: the brute-force sum of the half spectrum is exact
    np.testing.assert_allclose(oracle.c2r(c, 8), brute_c2r(c, 8) * 0 + x * x.size, atol=1e-9)
    np.testing.assert_allclose(brute_c2r(c, 8), x * x.size, atol=1e-9)


def test_roundtrip_scales_by_n3(oracle):
    rng = np.random.default_rng(5)
    x = rng.standard_normal((8, 16, 32))
    np.testing.assert_allclose(oracle.c2r(oracle.r2c(x), 32), x * x.size, rtol=1e-12, atol=1e-10)


def test_single_mode_known_answer(oracle):
    n = 16
    g = np.arange(n) * 2 * np.pi / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    f = np.cos(x) + 2 * np.sin(3 * y) - 0.5 * np.cos(2 * z + x)
    c = oracle.r2c(f)
    N3 = n**3
    assert np.isclose(c[0, 0, 1], N3 / 2)
    assert np.isclose(c[0, 3, 0], -1j * N3)  # 2 sin(3y) = -i e^{3iy} + i e^{-3iy}
    assert np.isclose(c[0, -3, 0], 1j * N3)
    assert np.isclose(c[2, 0, 1], -0.25 * N3)
    mask = np.ones_like(c, bool)
    for idx in [(0, 0, 1), (0, 3, 0), (0, -3, 0), (2, 0, 1)]:
        mask[idx] = False
    assert np.abs(c[mask]).max() < 1e-9 * N3
    # c2r puts the physical field back (times N^3)
    np.testing.assert_allclose(oracle.c2r(c, n), f * N3, atol=1e-9 * N3)


def test_c2r_ignores_imag_of_dc_and_nyquist(oracle):
    rng = np.random.default_rng(6)
    c = oracle.r2c(rng.standard_normal((4, 4, 8)))
    d = c.copy()
    d[0, 0, 0] += 3j
    d[0, 0, 4] -= 2j
    np.testing.assert_allclose(oracle.c2r(c, 8), oracle.c2r(d, 8), atol=1e-12)


def test_parseval(oracle):
    rng = np.random.default_rng(7)
    x = rng.standard_normal((16, 16, 16))
    c = oracle.r2c(x)
    w = np.ones(c.shape)
    w[:, :, 1:-1] = 2.0
    assert np.isclose((w * np.abs(c) ** 2).sum() / x.size, (x**2).sum())
