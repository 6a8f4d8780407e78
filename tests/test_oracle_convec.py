"""Pins the oracle's vorticity (dns-deform.f90:265-273) and convec_dns
(convec_dns.f90:15-53) by closed forms (ABC flow is Beltrami, so u x w = 0 and
w = u) and by brute-force physical-space sums on an 8^3 grid."""
import numpy as np


def _grid(n):
    g = np.arange(n) * 2 * np.pi / n
    return np.meshgrid(g, g, g, indexing="ij")  # z, y, x


def test_abc_flow_is_beltrami(oracle):
    n = 16
    z, y, x = _grid(n)
    A, B, C = 1.0, 0.7, 0.4
    ux = A * np.sin(z) + C * np.cos(y)
    uy = B * np.sin(x) + A * np.cos(z)
    uz = C * np.sin(y) + B * np.cos(x)
    kx, ky, kz, k2 = oracle.wavenumber(n, n, n)
    v = [oracle.r2c(a) / n**3 for a in (ux, uy, uz)]
    w = oracle.vorticity(*v, kx, ky, kz)
    for a, b in zip(v, w):  # curl u = u for the k=1 ABC flow
        np.testing.assert_allclose(a, b, atol=1e-14)
    out = oracle.convec_dns(*v, *w, n, n, n)
    assert max(np.abs(a).max() for a in out) < 1e-15


def test_taylor_green_closed_form(oracle):
    n = 16
    z, y, x = _grid(n)
    ux = np.sin(x) * np.cos(y) * np.cos(z)
    uy = -np.cos(x) * np.sin(y) * np.cos(z)
    uz = np.zeros_like(x)
    # w = curl u (by hand)
    wx = -np.cos(x) * np.sin(y) * np.sin(z)
    wy = -np.sin(x) * np.cos(y) * np.sin(z)
    wz = 2 * np.sin(x) * np.sin(y) * np.cos(z)
    lamb = np.stack([uy * wz - uz * wy, uz * wx - ux * wz, ux * wy - uy * wx])
    kx, ky, kz, k2 = oracle.wavenumber(n, n, n)
    v = [oracle.r2c(a) / n**3 for a in (ux, uy, uz)]
    w = oracle.vorticity(*v, kx, ky, kz)
    for a, b in zip(w, (wx, wy, wz)):
        np.testing.assert_allclose(oracle.c2r(a, n), b, atol=1e-13)
    out = oracle.convec_dns(*v, *w, n, n, n)
    for a, b in zip(out, lamb):
        np.testing.assert_allclose(a, oracle.r2c(b) / n**3, atol=1e-14)


def test_convec_brute_force_8(oracle):
    """Direct Fourier-series sums (no FFT) for u, w and the DFT of u x w."""
    n = 8
    rng = np.random.default_rng(11)
    kx, ky, kz, k2 = oracle.wavenumber(n, n, n)
    v = []
    for _ in range(3):
        c = rng.standard_normal(k2.shape) + 1j * rng.standard_normal(k2.shape)
        c[k2 >= 9] = 0  # band-limited, so the products are alias-free
        oracle.symmetrize3d(c)
        v.append(c)
    oracle.projection(*v, kx, ky, kz)
    w = oracle.vorticity(*v, kx, ky, kz)
    out = oracle.convec_dns(*v, *w, n, n, n)

    z, y, x = _grid(n)
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)  # (n^3, 3)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    ks = np.stack([KX.ravel(), KY.ravel(), KZ.ravel()], 1)
    wgt = np.where(ks[:, 0] == 0, 1.0, 2.0)  # half spectrum: kx>0 twice
    E = np.exp(1j * pts @ ks.T)  # (n^3, modes)

    def phys(c):
        return (E * (wgt * c.ravel())[None, :]).real.sum(1)

    def dphys(c, j):  # derivative along axis j by the series definition
        return (E * (wgt * 1j * ks[:, j] * c.ravel())[None, :]).real.sum(1)

    u = [phys(c) for c in v]
    # vorticity from derivatives of the series (curl by definition)
    om = [dphys(v[2], 1) - dphys(v[1], 2), dphys(v[0], 2) - dphys(v[2], 0), dphys(v[1], 0) - dphys(v[0], 1)]
    lamb = [u[1] * om[2] - u[2] * om[1], u[2] * om[0] - u[0] * om[2], u[0] * om[1] - u[1] * om[0]]
    for a, b in zip(out, lamb):
        spec = (np.conj(E) * b[:, None]).sum(0) / n**3
        np.testing.assert_allclose(a.ravel(), spec, atol=1e-12)
