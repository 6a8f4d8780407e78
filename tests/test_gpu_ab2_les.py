"""GPU parity of the IF-AB2 stepper on the LES + scalar solver (the loops of
spectral_les.f90:241-304 and spectral_dns_scalar.f90:257-338: CFL dt on
the device, Euler and leapfrog bootstrap, the scalar with exp(-nu/Sc k^2 t))
against oracle/ab2.py driving the LES + scalar RHS of oracle/les_scalar.py.

Bars: FP64 velocity and scalar within 1e-10 after 12 CFL steps (static and
dynamic Smagorinsky, the dynamic coefficient refreshed every 3 steps) and
the dt sequence within 1e-13; FP32 at 64^3 within 2e-5 of the FP64 oracle
after 15 steps, graph replay bitwise equal to eager steps."""
import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import ab2
from oracle import les_scalar as les
from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu

NU, SC = 5e-3, 0.7
CFL = dict(beta=0.5, dtw=0.03, ieout=3, umax=3.0 * np.sqrt(3.0))


def relerr(a, b):
    return np.abs(np.asarray(a, np.complex128) - b).max() / np.abs(b).max()


def _fields(n):
    v = orc.initialize(n, n, n, 1, 4.0, 1.0)
    phi = orc.initialize(n, n, n, -10, 3.0, 1.0)[0]
    return v, phi


def _oracle(n, dynamic, v, phi, iupdate=3):
    delta = 2 * np.pi / n
    prm = les.LESParams(nx=n, nu=NU, dt=0.0, forced=True, fkmax=2.5, eps=0.1, cs2=0.0289, delta=delta,
                        dynamic=dynamic, iupdate=iupdate, scalar=True, G=1.0, Sc=SC)
    o = les.LESScalar(prm)
    o.rhs_per_step = 1  # one RHS a step: the dynamic update every iupdate steps
    a = ab2.AB2(ab2.AB2Params(nx=n, nu=NU, forced=True, fkmax=2.5, eps=0.1, kcut=n // 2 - 1.0, **CFL),
                [x.astype(np.complex128) for x in (*v, phi)], rhs=o.rhs, nus=[NU, NU, NU, NU / SC])
    return a, o


def _solver(n, prec, dynamic, use_graph=False, iupdate=3):
    delta = 2 * np.pi / n
    kc = 2 * (n // 2) / 3
    p = scl.Plan3D(n, n, n, prec)
    s = scl.DNS(p, nu=NU, forced=1, fkmax=2.5, eps=0.1, sgs=2 if dynamic else 1, cs2=0.0289, delta=delta,
                delta_test=2 * delta, k_test=kc / 2, iupdate=iupdate, scalar=True, grad=1.0, schmidt=SC,
                stepper=1, kcut=n // 2 - 1.0, use_graph=use_graph, **CFL)
    return p, s


@pytest.mark.parametrize("dynamic", [False, True])
def test_les_scalar_ab2_fp64(dynamic):
    n, nsteps = 32, 12
    print(f"case ab2 les+scalar fp64 dynamic={dynamic}", flush=True)
    v, phi = _fields(n)
    o, lo = _oracle(n, dynamic, v, phi)
    p, s = _solver(n, scl.DP, dynamic)
    s.set_velocity(*[a.copy() for a in v])
    s.set_scalar(phi.copy())
    dts = []
    for _ in range(nsteps):
        s.step(1)
        o.step()
        dts.append(s.step_info()[1])
    derr = np.abs(np.array(dts) - np.array(o.dts)).max() / max(o.dts)
    ev = max(relerr(a, b) for a, b in zip(s.get_velocity(), o.v[:3]))
    ep = relerr(s.get_scalar(), o.v[3])
    print(f"velocity {ev:.2e}, scalar {ep:.2e}, dt {derr:.2e}, dts {np.round(dts, 6)}", flush=True)
    assert derr < 1e-13 and ev < 1e-10 and ep < 1e-10
    assert np.std(dts) > 0.05 * np.mean(dts)
    if dynamic:
        assert np.isclose(s.diagnostics()["cs2"], lo.cs2, rtol=1e-9)
    s.close()
    p.close()


def test_les_scalar_ab2_fp32_and_graph():
    n, nsteps = 64, 15
    print("case ab2 les+scalar fp32 64^3, graph vs eager", flush=True)
    v, phi = _fields(n)
    o, _ = _oracle(n, False, v, phi)
    for _ in range(nsteps):
        o.step()
    out = []
    for g in (True, False):
        p, s = _solver(n, scl.SP, False, use_graph=g)
        s.set_velocity(*[a.astype(np.complex64) for a in v])
        s.set_scalar(phi.astype(np.complex64))
        s.step(4)
        s.step(nsteps - 4)
        out.append((*s.get_velocity(), s.get_scalar()))
        s.close()
        p.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
    err = max(relerr(a, b) for a, b in zip(out[0], o.v))
    print(f"fp32 15 steps rel err {err:.2e}", flush=True)
    assert err < 2e-5
