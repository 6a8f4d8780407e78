"""Pins of oracle/turb2d.py (turb_2d.f90 at one rank).

* the transform pair against the Ooura definitions quoted in
  fft1d_mpi.f90:17-67 (direct cos/sin and exp sums, with newfft's scalings
  and a(1) zeroing, turb_2d.f90:660-723);
* the phase-shift dealiased nonlinear term against the exact convolution
  -i k.F(u w), computed on a 4x finer grid, for vorticity band-limited to
  the 8/9 disk (the alias-free case of the scheme); without the shifted
  half the same fields alias;
* the stepper: with the nonlinear term off and one mode of known decay rate
  (sigma = 1/5, nu = 5 at |k| = 5: L = 2), the closed form
  w0 (1 - L dt + (L dt)^2/2) exp(-L dt)^(m-2) and the time sequence
  dt/2, dt/2, dt...;  with a constant nonlinear term c and forcing,
  first-order convergence to dw/dt = -L w + c + (3 f(w) - f(c))/2, the
  limit of the update as written (old forcing evaluated at owz = c).
Each check also fails on a copy with a planted mistake."""
import inspect

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle import turb2d as t2


def _mutant(fn, old, new, *more):
    src = inspect.getsource(fn)
    for a, b in ((old, new),) + tuple(zip(more[::2], more[1::2])):
        assert a in src, a
        src = src.replace(a, b)
    ns = dict(inspect.getmodule(fn).__dict__)
    exec(compile(src, "<mutant>", "exec"), ns)
    return ns[fn.__name__]


def _ooura_to_x(c):
    """cdft(-1) along y, a(1) = 0, rdft(-1) along x, as the quoted sums."""
    n = c.shape[0]
    y = np.arange(n)
    out = np.zeros((n, n))
    chat = np.zeros((n, n // 2), complex)
    for ix in range(n // 2):
        for yy in range(n):
            chat[yy, ix] = np.sum(c[:, ix] * np.exp(-2j * np.pi * y * yy / n))
    chat[:, 0] = chat[:, 0].real
    for yy in range(n):
        for k in range(n):
            a = chat[yy, 0].real / 2.0
            for j in range(1, n // 2):
                a += chat[yy, j].real * np.cos(2 * np.pi * j * k / n) + chat[yy, j].imag * np.sin(2 * np.pi * j * k / n)
            out[yy, k] = a
    return out


def _ooura_to_k(u):
    """rdft(+1) along x, / (n/2), a(1) = 0; cdft(+1) along y, / n."""
    n = u.shape[0]
    x = np.arange(n)
    r = np.zeros((n, n // 2), complex)
    for yy in range(n):
        for k in range(n // 2):
            re = np.sum(u[yy] * np.cos(2 * np.pi * x * k / n))
            im = np.sum(u[yy] * np.sin(2 * np.pi * x * k / n)) if k else 0.0
            r[yy, k] = (re + 1j * im) / (n / 2)
    out = np.zeros_like(r)
    for ix in range(n // 2):
        for k in range(n):
            out[k, ix] = np.sum(r[:, ix] * np.exp(2j * np.pi * x * k / n)) / n
    return out


def _check_transforms(to_x, to_k):
    n = 8
    rng = np.random.default_rng(1)
    c = rng.standard_normal((n, n // 2)) + 1j * rng.standard_normal((n, n // 2))
    c = t2.symmetrize(c)
    u = rng.standard_normal((n, n))
    assert np.abs(to_x(c) - _ooura_to_x(c)).max() < 1e-12
    assert np.abs(to_k(u) - _ooura_to_k(u)).max() < 1e-12


def test_transforms_follow_the_ooura_definitions():
    _check_transforms(t2.to_x, t2.to_k)


@pytest.mark.parametrize("fn,old,new", [("to_x", "full[:, 0] *= 0.5", "pass"),
                                        ("to_x", "np.fft.fft2(full)", "np.fft.ifft2(full) * full.size"),
                                        ("to_k", "2.0 * np.fft.ifft2(u)", "np.fft.ifft2(u)")])
def test_transform_check_catches_planted_mistakes(fn, old, new):
    f = {"to_x": t2.to_x, "to_k": t2.to_k}
    f[fn] = _mutant(f[fn], old, new)
    with pytest.raises(AssertionError):
        _check_transforms(f["to_x"], f["to_k"])


def _band_limited(n, seed):
    kx, ky, k2 = t2.kgrid(n)
    ass = float(np.float32(np.float32(2.0) / np.float32(9.0)) * np.float32(n * n))
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n, n // 2)) + 1j * rng.standard_normal((n, n // 2))
    w = np.where(k2 < ass, w, 0)
    return t2.symmetrize(w), kx, ky, k2, ass


def _exact_nonlinear(w, kx, ky, k2, ass):
    """-i (kx F(u w) + ky F(v w)) from products on a grid 4x finer."""
    n = w.shape[0]
    m = 4 * n

    def fine(c):
        f = np.zeros((m, m // 2), complex)
        iy = (ky[:, 0].astype(int)) % m
        f[iy, : n // 2] = c
        return t2.to_x(f)

    ww = fine(w)
    pu = t2.to_k(fine(1j * ky * w / k2) * ww)
    pv = t2.to_k(fine(-1j * kx * w / k2) * ww)
    iy = (ky[:, 0].astype(int)) % m
    ex = -1j * (kx * pu[iy, : n // 2] + ky * pv[iy, : n // 2])
    ex = np.where(k2 < ass, ex, 0)
    ex[0, 0] = 0
    return ex


def _check_nonlinear(fn):
    n = 16
    w, kx, ky, k2, ass = _band_limited(n, 2)
    want = _exact_nonlinear(w, kx, ky, k2, ass)
    got = fn(w)
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < 1e-12, err
    return w, want


def test_nonlinear_is_the_exact_dealiased_convolution():
    w, want = _check_nonlinear(t2.nonlinear)
    # the unshifted half alone (doubled) aliases at this radius
    half = _mutant(t2.nonlinear, "vy = vy + c3 * (-0.5j) * kx", "vy = 2 * vy", "vy = vy + c4 * (-0.5j) * ky", "pass")
    alias = half(w)
    assert np.abs(alias - want).max() / np.abs(want).max() > 1e-3


@pytest.mark.parametrize("old,new", [("* np.conj(ph))\n    vy = vy + c3", "* ph)\n    vy = vy + c3"),
                                     ("wts = to_x(symmetrize(w * ph))", "wts = to_x(symmetrize(w))"),
                                     ("vy = a * (-0.5j) * kx", "vy = a * (0.5j) * kx"),
                                     ("ph = np.exp(1j * np.pi / n * (kx + ky))", "ph = np.exp(2j * np.pi / n * (kx + ky))")])
def test_nonlinear_check_catches_planted_mistakes(old, new):
    with pytest.raises(AssertionError):
        _check_nonlinear(_mutant(t2.nonlinear, old, new))


def _check_linear(cls):
    n, dt, m = 16, 0.01, 12
    w0 = np.zeros((n, n // 2), complex)
    w0[4, 3] = 0.7 - 0.2j  # kx = 3, ky = 4: |k| = 5
    p = t2.Turb2DParams(n=n, dt=dt, sigma=0.2, nu=5.0, eta=1.0, kf=1000.0)
    s = cls(p, w0, nonlinear_fn=lambda w: np.zeros_like(w))
    times = []
    for _ in range(m):
        s.step()
        times.append(s.time)
    L = 2.0
    want = w0[4, 3] * (1 - L * dt + (L * dt) ** 2 / 2) * np.exp(-L * dt) ** (m - 2)
    assert abs(s.w[4, 3] - want) < 1e-14 * abs(want)
    assert np.count_nonzero(s.w) == 1
    np.testing.assert_allclose(times, [dt / 2, dt] + [dt * (k + 1) for k in range(1, m - 1)], rtol=0, atol=1e-15)


def test_stepper_linear_closed_form():
    _check_linear(t2.Turb2D)


def _check_forcing(cls):
    n, T = 16, 0.4
    rng = np.random.default_rng(3)
    p = t2.Turb2DParams(n=n, dt=0.0, sigma=0.01, nu=10.0, eta=60.0, kf=4.0)
    w0 = t2.symmetrize(rng.standard_normal((n, n // 2)) + 1j * rng.standard_normal((n, n // 2)) + 0.5)
    c = t2.symmetrize(0.3 * (rng.standard_normal((n, n // 2)) + 1j * rng.standard_normal((n, n // 2))) + 0.2)
    probe = cls(p, w0, nonlinear_fn=lambda w: c)
    band, L, etak = probe.band, probe.L, probe.etak
    (iy, ix) = [a[0] for a in np.nonzero(band & (np.abs(w0) > 0.3) & (np.abs(c) > 0.1))]
    lam, cc, wi = L[iy, ix], c[iy, ix], w0[iy, ix]
    f = lambda a: etak / np.conj(a)  # noqa: E731

    def rhs(t, y):
        w = y[0] + 1j * y[1]
        d = -lam * w + cc + (3 * f(w) - f(cc)) / 2
        return [d.real, d.imag]

    errs = []
    for nsteps in (400, 800):
        p.dt = T / nsteps
        s = cls(p, w0, nonlinear_fn=lambda w: c)
        for _ in range(nsteps):  # the bootstrap's two steps reach dt
            s.step()
        ref = _at(rhs, wi, s.time)
        errs.append(abs(s.w[iy, ix] - ref))
    ratio = errs[1] / errs[0]
    assert errs[0] < 2e-2 * abs(ref) and 0.4 < ratio < 0.6, (errs, ratio)


def _at(rhs, wi, t):
    y = solve_ivp(rhs, (0, t), [wi.real, wi.imag], rtol=1e-12, atol=1e-14).y[:, -1]
    return y[0] + 1j * y[1]


def test_stepper_forcing_converges_to_its_limit():
    _check_forcing(t2.Turb2D)


@pytest.mark.parametrize("check,old,new", [
    ("linear", "self.w = w * (1.0 - L * dth)", "self.w = w * (1.0 - L * dt)"),
    ("linear", "E = np.where(self.k2 > 0.9, np.exp(-L * dt), 1.0)", "E = np.where(self.k2 > 0.9, np.exp(-L * dth), 1.0)"),
    ("linear", "self.w = self.w0 + dt * (n + self.f(w)) - L * dt * w", "self.w = w + dt * (n + self.f(w)) - L * dt * w"),
    ("forcing", "E * (self.owz + self.f(self.owz))", "E * (self.owz + self.f(w))"),
    ("forcing", "3.0 * (n + self.f(w))", "3.0 * (n + 0.5 * self.f(w))"),
])
def test_stepper_checks_catch_planted_mistakes(check, old, new):
    src = inspect.getsource(t2.Turb2D)
    assert old in src
    ns = dict(t2.__dict__)
    exec(compile(src.replace(old, new), "<mutant>", "exec"), ns)
    with pytest.raises(AssertionError):
        (_check_linear if check == "linear" else _check_forcing)(ns["Turb2D"])
