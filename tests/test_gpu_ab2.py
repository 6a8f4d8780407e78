"""GPU parity of the IF-AB2 steppers (spectral_dns_dealiased.f90:124-296
with CFL steps; specdns-forced-omp-ft.f90:224-274 with fixed dt) against
oracle/ab2.py, through the solver's C ABI.

Bars: FP64 fields within 1e-11 of the FP64 oracle after 14 steps and the dt
sequence (umax refreshes, landing on output times) within 1e-14; FP32
fields within 2e-5 after 20 steps at 64^3; an ABC flow decays exactly
(1e-13, FP64); graph replay bitwise equal to eager steps."""
import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import ab2
from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return max(np.abs(np.asarray(x, np.complex128) - y).max() for x, y in zip(a, b)) / max(np.abs(y).max() for y in b)


CFL = dict(nu=0.01, forced=True, fkmax=2.0, eps=0.1, beta=0.5, dtw=0.025, ieout=3, umax=3.0 * np.sqrt(3.0))


@pytest.mark.parametrize("use_graph", [False, True])
def test_cfl_stepper_fp64_matches_oracle(use_graph):
    n, nsteps = 32, 14
    print(f"case ab2 cfl fp64 graph={use_graph}", flush=True)
    o = ab2.AB2(ab2.AB2Params(nx=n, kcut=n // 2 - 1.0, **CFL), orc.initialize(n, n, n, 1, 4.0, 1.0))
    p = scl.Plan3D(n, n, n, scl.DP)
    s = scl.DNS(p, nu=CFL["nu"], forced=1, fkmax=2.0, eps=0.1, stepper=1, beta=CFL["beta"], dtw=CFL["dtw"],
                ieout=CFL["ieout"], umax=CFL["umax"], kcut=n // 2 - 1.0, use_graph=use_graph)
    s.set_velocity(*[a.copy() for a in orc.initialize(n, n, n, 1, 4.0, 1.0)])
    dts = []
    for _ in range(nsteps):
        s.step(1)
        o.step()
        dts.append(s.step_info()[1])
    t, _, k = s.step_info()
    err = relerr(s.get_velocity(), o.v)
    derr = np.abs(np.array(dts) - np.array(o.dts)).max() / max(o.dts)
    print(f"fields rel err {err:.2e}, dt rel err {derr:.2e}, t {t} vs {o.time}, dts {np.round(dts, 6)}", flush=True)
    assert k == nsteps and abs(t - o.time) < 1e-14
    assert derr < 1e-14 and err < 1e-11
    assert np.std(dts) > 0.1 * np.mean(dts)  # CFL steps, landings and umax refreshes varied dt
    s.close()
    p.close()


def test_fixed_dt_stepper_fp64_matches_oracle():
    n, nsteps, dt = 32, 10, 0.004
    print("case ab2 fixed dt fp64", flush=True)
    kc = 2 * (n // 2) / 3
    o = ab2.AB2(ab2.AB2Params(nx=n, nu=0.01, forced=True, fkmax=2.0, eps=0.1, dt_fixed=dt, kcut=kc),
                orc.initialize(n, n, n, 2 - 5, 4.0, 1.0))
    p = scl.Plan3D(n, n, n, scl.DP)
    s = scl.DNS(p, nu=0.01, dt=dt, forced=1, fkmax=2.0, eps=0.1, stepper=2)
    s.set_velocity(*[a.copy() for a in orc.initialize(n, n, n, 2 - 5, 4.0, 1.0)])
    s.step(nsteps)
    for _ in range(nsteps):
        o.step()
    t, _, _ = s.step_info()
    err = relerr(s.get_velocity(), o.v)
    print(f"fixed dt: fields rel err {err:.2e}, t {t} vs {o.time}", flush=True)
    assert abs(t - o.time) < 1e-14 and err < 1e-11
    s.close()
    p.close()


def test_abc_flow_decays_exactly():
    n, nu = 32, 0.05
    print("case ab2 abc decay", flush=True)
    x = 2 * np.pi * np.arange(n) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    v0 = [orc.r2c(a) / n**3 for a in (np.sin(Z) + np.cos(Y), np.sin(X) + np.cos(Z), np.sin(Y) + np.cos(X))]
    p = scl.Plan3D(n, n, n, scl.DP)
    s = scl.DNS(p, nu=nu, forced=0, stepper=1, beta=0.3, dtw=0.05, ieout=2, umax=3.0)
    s.set_velocity(*[a.copy() for a in v0])
    s.step(25)
    t = s.time
    f = np.exp(-nu * t)
    err = relerr(s.get_velocity(), [f * a for a in v0])
    print(f"abc decay to t={t:.4f}: rel err {err:.2e}", flush=True)
    assert err < 1e-13
    s.close()
    p.close()


def test_fp32_run_and_graph_bitwise():
    n, nsteps = 64, 20
    print("case ab2 fp32 64^3 vs fp64 oracle, graph vs eager", flush=True)
    v0 = scl.initialize(n, n, n, 1, 4.0, 1.0, scl.SP)
    o = ab2.AB2(ab2.AB2Params(nx=n, kcut=n // 2 - 1.0, **CFL), [a.astype(np.complex128) for a in v0])
    for _ in range(nsteps):
        o.step()
    p = scl.Plan3D(n, n, n, scl.SP)
    out = []
    for g in (True, False):
        s = scl.DNS(p, nu=CFL["nu"], forced=1, fkmax=2.0, eps=0.1, stepper=1, beta=CFL["beta"], dtw=CFL["dtw"],
                    ieout=CFL["ieout"], umax=CFL["umax"], kcut=n // 2 - 1.0, use_graph=g)
        s.set_velocity(*v0)
        s.step(7)
        s.step(nsteps - 7)
        s.sync()
        out.append(s.get_velocity())
        s.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
    err = relerr(out[0], o.v)
    print(f"fp32 20 steps rel err {err:.2e}", flush=True)
    assert err < 2e-5
    p.close()


def test_heli_dns_loop_fp64():
    """spectral_heli_dns.f90's loop: the initialize_heli start, force_heli in
    the RHS, IF-AB2 with CFL steps; FP64 against the oracle (1e-11)."""
    n, nsteps, eps, eta = 32, 12, 0.1, 0.05
    print("case ab2 helical dns fp64", flush=True)
    v0 = orc.initialize_heli(n, n, n, -3, 0.8, 4.0, 1.0)
    d = orc.DNS(orc.DNSParams(nx=n, nu=CFL["nu"], dt=0.0, forced=False))

    def rhs(vx, vy, vz):
        om = orc.vorticity(vx, vy, vz, d.kx, d.ky, d.kz)
        nl = orc.convec_dns(vx, vy, vz, *om, n, n, n)
        f = orc.force_heli(vx, vy, vz, *om, d.k2, 2.0, eps, eta)
        w = [a + b for a, b in zip(nl, f)]
        for a in w:
            orc.symmetrize3d(a)
        orc.projection(*w, d.kx, d.ky, d.kz)
        return list(w)

    kw = {k: v for k, v in CFL.items() if k not in ("forced", "fkmax", "eps")}
    o = ab2.AB2(ab2.AB2Params(nx=n, kcut=n // 2 - 1.0, forced=False, **kw), v0, rhs=rhs)
    p = scl.Plan3D(n, n, n, scl.DP)
    s = scl.DNS(p, forced=2, fkmax=2.0, eps=eps, eta=eta, stepper=1, kcut=n // 2 - 1.0, **kw)
    s.initialize_heli(-3, 0.8, 4.0, 1.0)
    for _ in range(nsteps):
        s.step(1)
        o.step()
    err = relerr(s.get_velocity(), o.v)
    print(f"helical ab2 rel err {err:.2e}, t {s.time} vs {o.time}", flush=True)
    assert err < 1e-11 and abs(s.time - o.time) < 1e-14
    s.close()
    p.close()
