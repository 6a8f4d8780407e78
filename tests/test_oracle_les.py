"""Pins the oracle's Smagorinsky stress (dynsmag.f90, smag.f90), the dynamic
coefficient, convecphi_dns and the LES + scalar step by brute-force series
sums on small grids, the SGS energy identity and closed forms."""
import numpy as np

from oracle import les_scalar as les
from oracle import spectral3d as s3


def _band_field(n, seed, kmax2=9):
    rng = np.random.default_rng(seed)
    kx, ky, kz, k2 = s3.wavenumber(n, n, n)
    v = []
    for _ in range(3):
        c = rng.standard_normal(k2.shape) + 1j * rng.standard_normal(k2.shape)
        c[k2 >= kmax2] = 0
        s3.symmetrize3d(c)
        v.append(c)
    s3.projection(*v, kx, ky, kz)
    return v, (kx, ky, kz, k2)


def _series(n, kx, ky, kz):
    g = np.arange(n) * 2 * np.pi / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    ks = np.stack([KX.ravel(), KY.ravel(), KZ.ravel()], 1)
    wgt = np.where(ks[:, 0] == 0, 1.0, 2.0)
    E = np.exp(1j * pts @ ks.T)

    def phys(c, j=None):
        f = c.ravel() if j is None else 1j * ks[:, j] * c.ravel()
        return (E * (wgt * f)[None, :]).real.sum(1)

    def dft(b):
        return (np.conj(E) * b[:, None]).sum(0) / n**3

    return phys, dft


def test_smag_tau_brute_force_8():
    n = 8
    v, (kx, ky, kz, k2) = _band_field(n, 1)
    cs2, delta = 0.0289, 2 * np.pi / n
    tau = les.smag_tau(*v, kx, ky, kz, cs2, delta, (n, n, n))
    phys, dft = _series(n, kx, ky, kz)
    du = [[phys(v[i], j) for j in range(3)] for i in range(3)]  # du[i][j] = d u_i / d x_j
    S = {(i, j): 0.5 * (du[i][j] + du[j][i]) for i, j in les.PAIRS}
    a = np.sqrt(2 * sum((1 if i == j else 2) * S[(i, j)] ** 2 for i, j in les.PAIRS))
    for q, (i, j) in enumerate(les.PAIRS):
        np.testing.assert_allclose(tau[q].ravel(), dft(-2 * cs2 * delta**2 * a * S[(i, j)]), atol=1e-13)


def test_sgs_energy_identity():
    """Energy input of -i k_j tau_ij equals <tau_ij S_ij> = -cs2 D^2 <|S|^3>."""
    n = 16
    v, (kx, ky, kz, k2) = _band_field(n, 2, kmax2=30)
    cs2, delta = 0.0289, 2 * np.pi / n
    t = les.smag_tau(*v, kx, ky, kz, cs2, delta, (n, n, n))
    K = (kx[None, None, :], ky[None, :, None], kz[:, None, None])
    f = [-1j * (K[0] * t[0] + K[1] * t[1] + K[2] * t[2]),
         -1j * (K[0] * t[1] + K[1] * t[3] + K[2] * t[4]),
         -1j * (K[0] * t[2] + K[1] * t[4] + K[2] * t[5])]
    inj = 2 * np.real(sum(np.conj(a) * b for a, b in zip(v, f)))
    inj = inj.sum() - 0.5 * inj[:, :, 0].sum()  # d/dt (1/2 <u^2>)
    Sp = [s3.c2r(a, n) for a in les.strain(*v, kx, ky, kz)]
    a = les._abs_strain(Sp)
    assert inj < 0
    assert np.isclose(inj, -cs2 * delta**2 * (a**3).mean(), rtol=1e-12)


def test_dynamic_coefficient_closed_forms():
    n = 16
    v, (kx, ky, kz, k2) = _band_field(n, 3, kmax2=30)
    delta = 2 * np.pi / n
    # test filter passing every mode: L = 0, so cs2 = 0 exactly
    g1 = np.ones_like(k2)
    tau, mm, lm, cs2 = les.dynsmag_update(*v, kx, ky, kz, g1, delta, 2 * delta, (n, n, n))
    assert mm > 0 and abs(lm) < 1e-12 * mm and abs(cs2) < 1e-12
    # a real test filter: cs2 is invariant to u -> a u and scales as 1/Delta^2
    g = les.g_test(k2, 2.5)
    _, _, _, c1 = les.dynsmag_update(*v, kx, ky, kz, g, delta, 2 * delta, (n, n, n))
    _, _, _, c2 = les.dynsmag_update(*[3.0 * a for a in v], kx, ky, kz, g, delta, 2 * delta, (n, n, n))
    _, _, _, c3 = les.dynsmag_update(*v, kx, ky, kz, g, 2 * delta, 4 * delta, (n, n, n))
    assert c1 != 0 and np.isclose(c1, c2, rtol=1e-10) and np.isclose(c3, c1 / 4, rtol=1e-10)
    # tau of the update is the static stress at the new coefficient
    t2 = les.smag_tau(*v, kx, ky, kz, c1, delta, (n, n, n))
    tau, *_ = les.dynsmag_update(*v, kx, ky, kz, g, delta, 2 * delta, (n, n, n))
    for a, b in zip(tau, t2):
        np.testing.assert_allclose(a, b, atol=1e-14)


def test_convecphi_brute_force_and_sign():
    n = 8
    # |k_i| <= 1 in each factor keeps u phi below the grid's Nyquist mode
    v, (kx, ky, kz, k2) = _band_field(n, 4, kmax2=4)
    rng = np.random.default_rng(5)
    phi = rng.standard_normal(k2.shape) + 1j * rng.standard_normal(k2.shape)
    phi[k2 >= 4] = 0
    s3.symmetrize3d(phi)
    out = les.convecphi_dns(*v, phi, kx, ky, kz, (n, n, n))
    phys, dft = _series(n, kx, ky, kz)
    p = phys(phi)
    u = [phys(a) for a in v]
    dp = [phys(phi, j) for j in range(3)]
    du = [phys(v[j], j) for j in range(3)]
    # +div(u phi) = u . grad phi + phi div u, by the series definition
    want = dft(sum(u[j] * dp[j] + p * du[j] for j in range(3)))
    np.testing.assert_allclose(out.ravel(), want, atol=1e-12)


def test_scalar_pure_diffusion_and_mean_gradient():
    n = 16
    p = les.LESParams(nx=n, nu=0.05, dt=0.01, forced=False, cs2=0.0, G=0.0, Sc=0.7)
    d = les.LESScalar(p)
    rng = np.random.default_rng(6)
    phi = rng.standard_normal(d.k2.shape) + 1j * rng.standard_normal(d.k2.shape)
    phi[d.k2 >= d.kcut2] = 0
    s3.symmetrize3d(phi)
    z = np.zeros_like(phi)
    out = (z, z, z, phi)
    for _ in range(5):
        out = d.rk4_step(*out)
    np.testing.assert_allclose(out[3], np.exp(-p.nu / p.Sc * d.k2 * 5 * p.dt) * phi, atol=1e-14)
    # phi = 0 and a uniform-in-time single velocity mode: dphi/dt = -G u_z to first order
    p2 = les.LESParams(nx=n, nu=0.0, dt=1e-4, forced=False, cs2=0.0, G=1.0)
    d2 = les.LESScalar(p2)
    u = [np.zeros_like(phi) for _ in range(3)]
    u[2][0, 1, 0] = 0.1  # u_z(y) = 0.2 cos y: a steady shear, no advection of phi = 0
    u[2][0, -1, 0] = 0.1
    o = d2.rk4_step(*u, np.zeros_like(phi))
    np.testing.assert_allclose(o[3][0, 1, 0], -p2.G * 0.1 * p2.dt, rtol=1e-9)
