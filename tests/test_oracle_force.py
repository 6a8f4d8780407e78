"""Pins the oracle's force (force.f90:16-29) and force_heli (force_heli.f90:15-55)
by their defining invariants: injected energy = eps, injected helicity = eta."""
import numpy as np


def _field(oracle, n, seed):
    v = oracle.initialize(n, n, n, seed, 4.0, 1.0)
    return [a.astype(np.complex128) for a in v]


def _hsum(t):
    return t.sum() - 0.5 * t[:, :, 0].sum()


def test_constant_power_injection(oracle):
    n = 16
    kx, ky, kz, k2 = oracle.wavenumber(n, n, n)
    v = _field(oracle, n, -3)
    w = [np.zeros_like(a) for a in v]
    a = oracle.force(*w, *v, k2, 2.0, 0.1)
    power = _hsum(2 * np.real(sum(np.conj(x) * f for x, f in zip(v, w))))
    assert np.isclose(power, 0.1, rtol=1e-12)
    ef = _hsum(np.where(k2 < 4, sum(np.abs(x) ** 2 for x in v), 0))
    assert np.isclose(a, 0.1 / (2 * ef), rtol=1e-12)
    assert all(np.all(f[k2 >= 4] == 0) for f in w)


def test_helical_injection(oracle):
    n = 16
    kx, ky, kz, k2 = oracle.wavenumber(n, n, n)
    v = _field(oracle, n, -5)
    w = oracle.vorticity(*v, kx, ky, kz)
    eps, eta = 0.1, 0.03
    f = oracle.force_heli(*v, *w, k2, 2.5, eps, eta)
    e_in = _hsum(2 * np.real(sum(np.conj(x) * g for x, g in zip(v, f))))
    assert np.isclose(e_in, eps, rtol=1e-10)
    curl_f = oracle.vorticity(*f, kx, ky, kz)
    # H = sum' (u.w* + u*.w); dH/dt from f = sum' 2 Re(f*.w + u*.curl f)
    h_in = _hsum(2 * np.real(sum(np.conj(g) * om + np.conj(x) * cf for g, om, x, cf in zip(f, w, v, curl_f))))
    assert np.isclose(h_in, eta, rtol=1e-9)
    assert all(g[0, 0, 0] == 0 for g in f)
    assert all(np.all(g[k2 >= 2.5**2] == 0) for g in f)
