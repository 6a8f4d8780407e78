"""Pins of the IF-AB2 oracle (oracle/ab2.py, spectral_dns_dealiased.f90:
124-296 and specdns-forced-omp-ft.f90:224-274):

* N = 0 (an unforced ABC flow, u x w = 0): every step multiplies each mode
  by exp(-nu k^2 dt) exactly, whatever the dt sequence (CFL steps, umax
  refreshes, landing on the output times), so v(t) = exp(-nu t) v0;
* a scalar linear ODE y' = -lam y + f(t) with a closed-form solution, on
  random variable steps: second-order convergence of the update formulas;
* the DNS at 32^3 from a k = 1 field to t = 0.02 (the cut far above the
  modes the run reaches): second-order convergence of the CFL-stepped run,
  and of the fixed-dt form, against a fine IF-RK4 run.

Each convergence check also fails on copies of the update with a planted
mistake (the dt/dto ratio dropped from a weight, the old RHS's integrating
factor dropped, the leapfrog weight of the bootstrap halved)."""
import inspect

import numpy as np
import pytest

from oracle import ab2
from oracle import spectral3d as orc


def _abc(n):
    x = 2 * np.pi * np.arange(n) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = [np.sin(Z) + np.cos(Y), np.sin(X) + np.cos(Z), np.sin(Y) + np.cos(X)]
    return [orc.r2c(a) / n**3 for a in u]


def test_exact_viscous_decay_with_variable_steps():
    n, nu = 16, 0.05
    p = ab2.AB2Params(nx=n, nu=nu, forced=False, beta=0.3, dtw=0.05, ieout=2, umax=3.0)
    v0 = _abc(n)
    s = ab2.AB2(p, v0)
    for _ in range(25):
        s.step()
    dts = np.array(s.dts)
    assert dts.std() > 0.05 * dts.mean()  # the steps did vary
    f = np.exp(-nu * s.time)
    err = max(np.abs(a - f * b).max() for a, b in zip(s.v, v0)) / max(np.abs(b).max() for b in v0)
    assert err < 1e-13, err
    # the output times were hit exactly: 0.05, 0.10, ...
    t = np.cumsum(dts)
    assert np.min(np.abs(t[:, None] - 0.05 * np.arange(1, 5)[None, :]), axis=0).max() < 1e-12


def _check_per_field_viscosity(cls):
    """With the RHS off, each field decays by its own exp(-nu_q k^2 t): the
    velocity with nu, the scalar with nu/Sc (spectral_dns_scalar.f90:266-322)."""
    n, nu, sc = 16, 0.05, 0.3
    p = ab2.AB2Params(nx=n, nu=nu, forced=False, beta=0.3, dtw=0.05, ieout=2, umax=3.0)
    rng = np.random.default_rng(4)
    v0 = _abc(n) + [orc.r2c(rng.standard_normal((n, n, n))) / n**3]
    for a in v0:
        a[orc.wavenumber(n, n, n)[3] >= (n // 2 - 1.0) ** 2] = 0
        a[0, 0, 0] = 0
    s = cls(p, v0, rhs=lambda *v: [np.zeros_like(a) for a in v], nus=[nu, nu, nu, nu / sc])
    for _ in range(12):
        s.step()
    k2 = orc.wavenumber(n, n, n)[3]
    for q, a in enumerate(s.v):
        want = v0[q] * np.exp(-(nu if q < 3 else nu / sc) * k2 * s.time)
        assert np.abs(a - want).max() < 1e-13 * np.abs(v0[q]).max(), q


def test_per_field_viscosity():
    _check_per_field_viscosity(ab2.AB2)


def test_per_field_viscosity_check_catches_shared_viscosity():
    src = inspect.getsource(ab2.AB2)
    old = "E_dt = np.exp(-self.nus[q] * self.k2 * dt)"
    assert old in src
    ns = dict(ab2.__dict__)
    exec(compile(src.replace(old, "E_dt = np.exp(-self.p.nu * self.k2 * dt)"), "<mutant>", "exec"), ns)
    with pytest.raises(AssertionError):
        _check_per_field_viscosity(ns["AB2"])


# --------------------------------------------------------------------------
def _scalar_run(update, lam, dts):
    """y' = -lam y + cos(t), y(0) = 1, through the stepper's update formulas
    (the RHS of a state at time t is cos(t): it does not depend on y)."""
    t, y, aux, dto, n0 = 0.0, 1.0, None, 1e30, None
    for i, dt in enumerate(dts):
        n = np.cos(t)
        E_dt = np.exp(-lam * dt)
        E_dto = np.exp(-lam * dto) if i > 0 else None
        y, a, dto_new = update(i, np.array(y), np.array(n), aux, dt, dto, E_dt, E_dto)
        if i == 0:
            aux, n0 = a, n
        elif i == 1:
            aux = np.array(n0)
        else:
            aux = a
        dto = dto_new
        t += dt
    return float(y), t


def _scalar_exact(lam, t):
    # y' = -lam y + cos t, y(0) = 1
    c = lam / (1 + lam * lam)
    d = 1 / (1 + lam * lam)
    return (1 - c) * np.exp(-lam * t) + c * np.cos(t) + d * np.sin(t)


def _check_scalar_order(update):
    rng = np.random.default_rng(5)
    base = rng.uniform(0.5, 1.5, 40)
    base *= 1.0 / base.sum()
    errs = []
    for lam in (0.7, 3.0):
        e = []
        for h in (1.0, 0.5, 0.25):
            dts = np.concatenate([base * h] * int(round(1 / h)))  # the same sequence, scaled, to t = 1
            y, t = _scalar_run(update, lam, dts)
            e.append(abs(y - _scalar_exact(lam, t)))
        errs.append(e)
        r1, r2 = e[0] / e[1], e[1] / e[2]
        assert 3.2 < r1 < 5.0 and 3.2 < r2 < 5.0, (lam, e, r1, r2)
    return errs


def test_scalar_ode_second_order():
    _check_scalar_order(ab2.ab2_update)


def _mutant(fn, old, new):
    src = inspect.getsource(fn)
    assert old in src, old
    ns = dict(inspect.getmodule(fn).__dict__)
    exec(compile(src.replace(old, new), "<mutant>", "exec"), ns)
    return ns[fn.__name__]


MUTANTS = [
    ("S = (1.0 + 0.5 * dtt) * dt", "S = 1.5 * dt"),
    ("ek = -0.5 * dtt * dt", "ek = -0.5 * dt"),
    ("ek * aux * E_dto", "ek * aux"),
    ("delta = 0.5 * dto * (1.0 + dtt) ** 2", "delta = 0.25 * dto * (1.0 + dtt) ** 2"),
    ("(S * aux * E_dto + ek * v", "(S * aux + ek * v"),
]


@pytest.mark.parametrize("old,new", MUTANTS)
def test_scalar_check_catches_planted_mistakes(old, new):
    bad = _mutant(ab2.ab2_update, old, new)
    with pytest.raises(AssertionError):
        _check_scalar_order(bad)


# --------------------------------------------------------------------------
T_END = 0.02


def _dns_ic(n):
    return orc.initialize(n, n, n, -3, 1.0, 1.0, kcut=1.5)


def _rk4_reference(n, nu):
    d = orc.DNS(orc.DNSParams(nx=n, nu=nu, dt=T_END / 80, forced=True, fkmax=2.0, eps=0.1))
    w = _dns_ic(n)
    for _ in range(80):
        w = d.rk4_step(*w)
    return w


def _ab2_run(n, nu, beta=None, dt_fixed=0.0):
    p = ab2.AB2Params(nx=n, nu=nu, forced=True, fkmax=2.0, eps=0.1, beta=beta or 0.1, dtw=T_END, ieout=1,
                      umax=3.0, kcut=15.0, dt_fixed=dt_fixed)
    s = ab2.AB2(p, _dns_ic(n))
    while s.time < T_END - 1e-12:
        s.step()
    assert abs(s.time - T_END) < 1e-12
    return s.v, len(s.dts)


def _check_dns_order(runs, ref):
    e = [max(np.abs(a - b).max() for a, b in zip(v, ref)) / max(np.abs(b).max() for b in ref) for v in runs]
    r1, r2 = e[0] / e[1], e[1] / e[2]
    assert 3.0 < r1 < 5.5 and 3.0 < r2 < 5.5, (e, r1, r2)


@pytest.fixture(scope="module")
def dns_ref():
    return _rk4_reference(32, 0.01)


def test_dns_cfl_run_second_order(dns_ref):
    runs = []
    for beta in (0.08, 0.04, 0.02):
        v, nsteps = _ab2_run(32, 0.01, beta=beta)
        runs.append(v)
    _check_dns_order(runs, dns_ref)


def test_dns_fixed_dt_second_order(dns_ref):
    runs = [_ab2_run(32, 0.01, dt_fixed=T_END / k)[0] for k in (6, 12, 24)]
    _check_dns_order(runs, dns_ref)


def test_dns_check_catches_dropped_ratio(dns_ref, monkeypatch):
    bad = _mutant(ab2.ab2_update, "ek = -0.5 * dtt * dt", "ek = -0.5 * dt")
    monkeypatch.setattr(ab2, "ab2_update", bad)
    runs = [_ab2_run(32, 0.01, beta=b)[0] for b in (0.08, 0.04, 0.02)]
    with pytest.raises(AssertionError):
        _check_dns_order(runs, dns_ref)
