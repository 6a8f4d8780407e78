"""Pins the oracle's 2D vorticity RHS and IF-RK4 (spectral_dns_2d.f90:94,
161-219 semantics): exact decay of a steady Euler flow, energy and enstrophy
conservation of the truncated nonlinear term, the forcing's enstrophy input,
and brute-force series sums on a 16^2 grid."""
import numpy as np

from oracle import vort2d as v2
from oracle import spectral3d as s3


def _field(n, seed, kmax2):
    rng = np.random.default_rng(seed)
    kx, ky, k2 = v2.wavenumber2d(n)
    w = rng.standard_normal(k2.shape) + 1j * rng.standard_normal(k2.shape)
    w[k2 >= kmax2] = 0
    s3.symmetrize2d(w)
    return w, k2


def test_steady_euler_mode_decays_by_the_integrating_factor():
    n = 32
    p = v2.Vort2DParams(n=n, dt=0.01, forced=False)
    d = v2.Vort2D(p)
    w = np.zeros((n, n // 2 + 1), complex)
    # modes with |k|^2 = 25: (3,4), (4,3), (5,0), (0,5): any combination is a steady Euler flow
    for kx, ky in [(3, 4), (4, -3), (5, 0), (0, 5)]:
        w[ky % n, kx] = 0.3 + 0.1j * kx
    s3.symmetrize2d(w)
    w0 = w.copy()
    for _ in range(4):
        w = d.rk4_step(w)
    np.testing.assert_allclose(w, np.exp(-d.rate * 4 * p.dt) * w0, atol=1e-15)


def test_nonlinear_term_conserves_energy_and_enstrophy():
    n = 32
    p = v2.Vort2DParams(n=n, forced=False)
    d = v2.Vort2D(p)
    w, k2 = _field(n, 1, d.kcut2)
    r = d.rhs(w)

    def hsum(t):
        return t.sum() - 0.5 * t[:, 0].sum()

    assert abs(hsum(2 * np.real(np.conj(w) * r))) < 1e-12 * hsum(np.abs(w) ** 2) * n
    assert abs(hsum(2 * np.real(np.conj(w) * r / k2))) < 1e-12 * hsum(np.abs(w) ** 2) * n


def test_forcing_enstrophy_input():
    n = 32
    p = v2.Vort2DParams(n=n, forced=True, eps=0.02)
    d = v2.Vort2D(p)
    w, k2 = _field(n, 2, d.kcut2)
    band = (k2 >= p.fkmin**2) & (k2 <= p.fkmax**2)
    pf = v2.Vort2DParams(n=n, forced=False)
    r_nf = v2.Vort2D(pf).rhs(w)
    f = d.rhs(w) - r_nf
    t = 2 * np.real(np.conj(w) * f)
    inj = t.sum() - 0.5 * t[:, 0].sum()
    cnt = band.sum() - 0.5 * band[:, 0].sum()
    assert np.isclose(inj, 2 * p.eps * cnt, rtol=1e-12)


def test_rhs_brute_force_16():
    n = 16
    p = v2.Vort2DParams(n=n, forced=False)
    d = v2.Vort2D(p)
    w, k2 = _field(n, 3, 9)
    r = d.rhs(w)
    g = np.arange(n) * 2 * np.pi / n
    y, x = np.meshgrid(g, g, indexing="ij")
    pts = np.stack([x.ravel(), y.ravel()], 1)
    KY, KX = np.meshgrid(d.ky, d.kx, indexing="ij")
    ks = np.stack([KX.ravel(), KY.ravel()], 1)
    wgt = np.where(ks[:, 0] == 0, 1.0, 2.0)
    E = np.exp(1j * pts @ ks.T)

    def phys(c):
        return (E * (wgt * c.ravel())[None, :]).real.sum(1)

    k2f = k2.ravel()
    om = phys(w)
    # velocity from the streamfunction psi = w/k^2: u = d psi/dy, v = -d psi/dx
    u = phys(1j * KY * w / k2)
    v = phys(-1j * KX * w / k2)
    dwx = phys(1j * KX * w)
    dwy = phys(1j * KY * w)
    adv = -(u * dwx + v * dwy)  # -u.grad w (= -div(u w), div u = 0)
    spec = (np.conj(E) * adv[:, None]).sum(0) / n**2
    spec = spec.reshape(k2.shape)
    spec[k2 >= d.kcut2] = 0
    s3.symmetrize2d(spec)
    np.testing.assert_allclose(r, spec, atol=1e-12)
