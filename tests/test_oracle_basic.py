"""Pins the oracle's wavenumber, symmetrize, dealiasing and projection
(wavenumber.f90:89-129, symmetrize.f90:32-64, dealiasing.f90:2-37,
projection.f90:2-31) by closed forms, index algebra and invariants."""
import numpy as np
import pytest


def test_wavenumber_closed_form(oracle):
    kx, ky, kz, k2 = oracle.wavenumber(8, 8, 4)
    assert kx.tolist() == [0, 1, 2, 3, 4]
    assert ky.tolist() == [0, 1, 2, 3, -4, -3, -2, -1]
    assert kz.tolist() == [0, 1, -2, -1]
    assert k2.shape == (4, 8, 5)
    assert k2[1, 2, 3] == 1 + 4 + 9
    assert k2[2, 4, 4] == 4 + 16 + 16
    assert k2[0, 0, 0] == np.finfo(np.float64).tiny
    _, _, _, k2s = oracle.wavenumber(8, 8, 8, "sp")
    assert k2s.dtype == np.float32 and k2s[0, 0, 0] == np.finfo(np.float32).tiny


def _full_spectrum(c, nx):
    """Hermitian-extend a half spectrum (nz, ny, nx/2+1) to (nz, ny, nx)."""
    nz, ny, lx1 = c.shape
    full = np.zeros((nz, ny, nx), complex)
    full[:, :, :lx1] = c
    for kx in range(lx1, nx):
        full[:, :, kx] = np.conj(c[(-np.arange(nz)) % nz][:, (-np.arange(ny)) % ny][:, :, nx - kx])
    return full


def test_symmetrize_makes_kx0_plane_hermitian(oracle):
    rng = np.random.default_rng(1)
    n = 8
    c = rng.standard_normal((n, n, n // 2 + 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1))
    orig = c.copy()
    oracle.symmetrize3d(c)
    # zeroed planes and mean
    assert np.all(c[:, n // 2, :] == 0) and np.all(c[n // 2, :, :] == 0)
    assert np.all(c[:, :, n // 2] == 0) and c[0, 0, 0] == 0
    # kx = 0 plane: Hermitian in (ky, kz)
    p = c[:, :, 0]
    pm = np.conj(p[(-np.arange(n)) % n][:, (-np.arange(n)) % n])
    np.testing.assert_array_equal(p, pm)
    # kept half untouched: kx in 1..nx/2-1 (off the zeroed planes), and
    # kx=0 with kz in 1..n/2-1
    keep = np.ones(c.shape, bool)
    keep[:, n // 2, :] = keep[n // 2, :, :] = False
    keep[:, :, n // 2] = False
    keep[:, :, 0] = False
    np.testing.assert_array_equal(c[keep], orig[keep])
    z = np.arange(1, n // 2)
    np.testing.assert_array_equal(c[z, :, 0][:, np.arange(n) != n // 2], orig[z, :, 0][:, np.arange(n) != n // 2])
    # the Hermitian-extended full spectrum inverts to a real field
    full = _full_spectrum(c, n)
    f = np.fft.ifftn(full)
    assert np.abs(f.imag).max() < 1e-14
    # idempotent
    d = c.copy()
    oracle.symmetrize3d(d)
    np.testing.assert_array_equal(c, d)


def test_symmetrize2d(oracle):
    rng = np.random.default_rng(2)
    n = 8
    c = rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))
    oracle.symmetrize2d(c)
    assert np.all(c[n // 2, :] == 0) and np.all(c[:, n // 2] == 0) and c[0, 0] == 0
    for y in range(1, n):
        assert c[y, 0] == np.conj(c[(n - y) % n, 0])


def test_dealiasing_truncates_at_ly2_over_9(oracle):
    rng = np.random.default_rng(3)
    n = 16
    kx, ky, kz, k2 = oracle.wavenumber(n, n, n)
    c = rng.standard_normal(k2.shape) + 1j * rng.standard_normal(k2.shape)
    oracle.dealiasing(c, k2)
    assert np.all(c[k2 >= n * n / 9] == 0)
    inside = (k2 < n * n / 9) & (np.abs(c) > 0)
    assert inside.sum() > 0.9 * ((k2 < n * n / 9).sum() - 3 * n * n)


def test_projection_invariants(oracle):
    rng = np.random.default_rng(4)
    n = 16
    kx, ky, kz, k2 = oracle.wavenumber(n, n, n)
    v = [rng.standard_normal(k2.shape) + 1j * rng.standard_normal(k2.shape) for _ in range(3)]
    KX, KY, KZ = kx[None, None, :], ky[None, :, None], kz[:, None, None]
    p = [a.copy() for a in v]
    oracle.projection(*p, kx, ky, kz)
    div = KX * p[0] + KY * p[1] + KZ * p[2]
    assert np.abs(div).max() < 1e-12 * np.abs(v[0]).max() * n
    q = [a.copy() for a in p]
    oracle.projection(*q, kx, ky, kz)
    for a, b in zip(p, q):
        np.testing.assert_allclose(a, b, atol=1e-13)
    # gradient fields are removed entirely
    phi = rng.standard_normal(k2.shape) + 1j * rng.standard_normal(k2.shape)
    g = [KX * phi, KY * phi, KZ * phi]
    oracle.projection(*g, kx, ky, kz)
    assert max(np.abs(a).max() for a in g) < 1e-12 * n
    # solenoidal part is untouched: P(v) + grad part = v with grad part || k
    for a, b in zip(v, p):
        pass
    r = [a - b for a, b in zip(v, p)]  # removed part is parallel to k
    cx = KY * r[2] - KZ * r[1]
    assert np.abs(cx).max() < 1e-12 * n * n
