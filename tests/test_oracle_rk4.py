"""Pins the oracle's IF-RK4 step and RHS (dns-deform.f90:160-237, 259-288):
exact viscous decay of a Beltrami flow (u x w = 0), the closed-form energy of
a forced Beltrami flow, and fourth-order convergence in dt."""
import numpy as np


def _abc(oracle, n, A=1.0, B=0.7, C=0.4):
    g = np.arange(n) * 2 * np.pi / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    u = (A * np.sin(z) + C * np.cos(y), B * np.sin(x) + A * np.cos(z), C * np.sin(y) + B * np.cos(x))
    return tuple(oracle.r2c(a) / n**3 for a in u)


def test_unforced_beltrami_decays_exactly(oracle):
    n = 16
    p = oracle.DNSParams(nx=n, nu=0.05, dt=0.01, forced=False)
    d = oracle.DNS(p)
    v0 = _abc(oracle, n)
    v = v0
    for _ in range(10):
        v = d.rk4_step(*v)
    fac = np.exp(-p.nu * 1.0 * 10 * p.dt)  # all modes have k^2 = 1
    for a, b in zip(v, v0):
        np.testing.assert_allclose(a, fac * b, atol=1e-14)


def test_forced_beltrami_energy_closed_form(oracle):
    n = 16
    nu, eps, dt, nsteps = 0.05, 0.2, 0.01, 50
    p = oracle.DNSParams(nx=n, nu=nu, dt=dt, forced=True, fkmax=2.0, eps=eps)
    d = oracle.DNS(p)
    v = _abc(oracle, n)
    e0 = oracle.energy(*v)
    for _ in range(nsteps):
        v = d.rk4_step(*v)
    t = nsteps * dt
    # dE/dt = eps - 2 nu E  (k = 1, and the forcing keeps the field Beltrami)
    e_exact = eps / (2 * nu) + (e0 - eps / (2 * nu)) * np.exp(-2 * nu * t)
    assert abs(oracle.energy(*v) - e_exact) < 1e-9 * e_exact


def test_fourth_order_convergence(oracle):
    n = 16
    T = 0.08
    v0 = oracle.initialize(n, n, n, 1, 3.0, 1.0)

    def run(dt):
        p = oracle.DNSParams(nx=n, nu=0.02, dt=dt, forced=True, fkmax=2.0, eps=0.1, dealias_rhs=True)
        d = oracle.DNS(p)
        v = v0
        for _ in range(int(round(T / dt))):
            v = d.rk4_step(*v)
        return np.concatenate([a.ravel() for a in v])

    ref = run(T / 64)
    errs = [np.abs(run(T / m) - ref).max() for m in (2, 4, 8)]
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(3.6 < r < 4.4 for r in rates), (errs, rates)


def test_rk4_keeps_field_solenoidal_and_truncated(oracle):
    n = 16
    p = oracle.DNSParams(nx=n)
    d = oracle.DNS(p)
    v = oracle.initialize(n, n, n, 1, 4.0, 1.0)
    for _ in range(3):
        v = d.rk4_step(*v)
    KX, KY, KZ = d.kx[None, None, :], d.ky[None, :, None], d.kz[:, None, None]
    div = np.abs(KX * v[0] + KY * v[1] + KZ * v[2]).max()
    assert div < 1e-14
    assert all(np.all(a[d.k2 >= p.kcut2] == 0) for a in v)
