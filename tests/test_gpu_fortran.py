"""The Fortran drop-in: a dns-deform-style program (tests/fortran/dns_deform_gpu.f90)
linked against the library's Fortran module set (reference routine names,
iso_c_binding over the C ABI) and run on the GPU, checked against the FP64
oracle.  Modes: 1 = RHS from convec_dns/force/symmetrize/projection with the
RK4 stages in Fortran; 2 = rk4step; 3 = as 1 with u x w built from the
rfftwnd_f77_one_* entry points; 4 = the LES + scalar RHS of
spectral_les.f90:205-238 / spectral_dns_scalar.f90 with dynsmag (coefficient
updated at the first RHS) and convecphi_dns by their reference names; 5 = the
LES with smag; 6 = scl_derivs bound to an FP64 solver, passed as the derivs
of a Fortran RK4 driver.  Bars: 2e-5 of max|v| after 5 steps at 32^3 (FP32
field from the library's FP32 initial condition); mode 6 (FP64) 1e-10."""
import os
import subprocess

import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import les_scalar as les
from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "speccodelib_amd", "fortran", "_build", "dns_deform_gpu")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (amdflang)")
@pytest.mark.parametrize("mode", [1, 2, 3])
def test_fortran_driver_matches_oracle(mode, tmp_path):
    n, nsteps = 32, 5
    out = tmp_path / f"v{mode}.bin"
    print(f"case fortran driver mode={mode}", flush=True)
    r = subprocess.run([EXE, str(n), str(nsteps), str(mode), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lx1 = n // 2 + 1
    v = np.fromfile(out, np.complex64).reshape(3, n, n, lx1)
    d = orc.DNS(orc.DNSParams(nx=n, nu=0.01, dt=0.002, forced=True, fkmax=2.0, eps=0.1))
    w = orc.initialize(n, n, n, 1, 4.0, 1.0)
    for _ in range(nsteps):
        w = d.rk4_step(*w)
    for a, b in zip(v, w):
        assert np.abs(a - b).max() < 2e-5 * np.abs(b).max()


def _run(mode, n, nsteps, out):
    r = subprocess.run([EXE, str(n), str(nsteps), str(mode), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout.strip(), flush=True)


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (amdflang)")
@pytest.mark.parametrize("mode", [4, 5])
def test_fortran_les_driver_matches_oracle(mode, tmp_path):
    n, nsteps = 32, 5
    out = tmp_path / f"les{mode}.bin"
    print(f"case fortran LES driver mode={mode}", flush=True)
    _run(mode, n, nsteps, out)
    lx1 = n // 2 + 1
    nf = 4 if mode == 4 else 3
    got = np.fromfile(out, np.complex64).reshape(nf, n, n, lx1)
    f32 = np.float32
    kcut = f32(2.0) * f32(n // 2) / f32(3)
    delta = float(f32(3.1415926536) / kcut)  # pi / kcut in the driver's precision
    cs2 = 0.0289 if mode == 4 else float(f32(0.026))  # smag.f90: c1 = 0.026 (real(sp))
    prm = les.LESParams(nx=n, nu=1e-3, dt=0.002, forced=True, fkmax=2.0, eps=0.1, cs2=cs2, delta=delta,
                        dynamic=mode == 4, iupdate=10**6, delta_ratio=2.0, scalar=mode == 4, G=1.0, Sc=0.7)
    o = les.LESScalar(prm)
    v = orc.initialize(n, n, n, 1, 4.0, 1.0)
    w = (*v, np.zeros_like(v[0]))
    for _ in range(nsteps):
        w = o.rk4_step(*w)
    print("oracle cs2", o.cs2, flush=True)
    for a, b in zip(got, w[:nf]):
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"rel err {err:.2e}", flush=True)
        assert err < 2e-5


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (amdflang)")
def test_fortran_derivs_driver_matches_oracle(tmp_path):
    # the driver's real(sp) parameters, widened to the FP64 solver's doubles
    f32 = np.float32
    n, nsteps, nu, dt, eps = 32, 5, float(f32(0.01)), float(f32(0.002)), float(f32(0.1))
    out = tmp_path / "derivs.bin"
    print("case fortran derivs driver", flush=True)
    _run(6, n, nsteps, out)
    lx1 = n // 2 + 1
    got = np.fromfile(out, np.complex128).reshape(3, n, n, lx1)
    d = orc.DNS(orc.DNSParams(nx=n, nu=nu, dt=dt, forced=True, fkmax=2.0, eps=eps))
    k2 = d.k2.astype(np.float64).copy()
    k2[0, 0, 0] = 0.0

    def f(y):
        return [a - nu * k2 * b for a, b in zip(d.rhs(*y), y)]

    # the driver's input: the library's FP32 initial field, widened
    y = [a.astype(np.complex128) for a in scl.initialize(n, n, n, 1, 4.0, 1.0, scl.SP)]
    for _ in range(nsteps):
        k1 = f(y)
        k2_ = f([a + 0.5 * dt * b for a, b in zip(y, k1)])
        k3 = f([a + 0.5 * dt * b for a, b in zip(y, k2_)])
        k4 = f([a + dt * b for a, b in zip(y, k3)])
        y = [a + dt / 6 * (b + 2 * c + 2 * e + g) for a, b, c, e, g in zip(y, k1, k2_, k3, k4)]
    err = max(np.abs(a - b).max() / np.abs(b).max() for a, b in zip(got, y))
    print(f"derivs RK4 rel err {err:.2e}", flush=True)
    assert err < 1e-10


PADDED_EXE = os.path.join(ROOT, "speccodelib_amd", "fortran", "_build", "padded_gpu")


@pytest.mark.skipif(not os.path.exists(PADDED_EXE), reason="Fortran driver not built (amdflang)")
def test_fortran_padded_routines(tmp_path):
    """padd, unpadd, the padded convec and smag_dealias.f90's smag called by
    their reference names from Fortran (libscl_fortran_padded.a linked before
    libscl_fortran.a), FP32, against oracle/padded.py on the 2n grid the
    library pads to (and padd/unpadd exactly)."""
    from oracle import padded as pd

    n = 32
    out = tmp_path / "padded.bin"
    print("case fortran padded routines", flush=True)
    r = subprocess.run([PADDED_EXE, str(n), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lx1, nb = n // 2 + 1, 3 * n // 2
    a = np.fromfile(out, np.complex64)
    m = n * n * lx1
    f = a[:16 * m].reshape(16, n, n, lx1)
    ub = a[16 * m:].reshape(nb, nb, nb // 2 + 1)
    v, w, lamb, tau, u2 = f[0:3], f[3:6], f[6:9], f[9:15], f[15]
    np.testing.assert_array_equal(ub, pd.padd(v[0], nb // 2 + 1, nb, nb))
    np.testing.assert_array_equal(u2, v[0])
    vd = [x.astype(np.complex128) for x in v]
    want = pd.convec_padded(*vd, *[x.astype(np.complex128) for x in w], (n, n, n), nb=(2 * n,) * 3)
    e1 = max(np.abs(x - y).max() for x, y in zip(lamb, want)) / max(np.abs(y).max() for y in want)
    kx, ky, kz, _ = orc.wavenumber(n, n, n)
    tw = pd.smag_dealias(*vd, kx, ky, kz, (n, n, n), float(np.float32(2 * np.float32(3.1415926536) / n)),
                         nb=(2 * n,) * 3)
    e2 = max(np.abs(x - y).max() for x, y in zip(tau, tw)) / max(np.abs(y).max() for y in tw)
    print(f"padded convec {e1:.2e}, smag_dealias {e2:.2e}", flush=True)
    assert e1 < 2e-6 and e2 < 2e-6


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (amdflang)")
def test_fortran_initialize_heli(tmp_path):
    """initialize_heli by its reference name (mode 7), FP32, against the
    FP64 oracle."""
    n = 16
    out = tmp_path / "heli.bin"
    print("case fortran initialize_heli", flush=True)
    _run(7, n, 0, out)
    v = np.fromfile(out, np.complex64).reshape(3, n, n, n // 2 + 1)
    want = orc.initialize_heli(n, n, n, -9, 0.6, 4.0, 1.0)
    err = max(np.abs(a - b).max() for a, b in zip(v, want)) / max(np.abs(b).max() for b in want)
    print(f"initialize_heli fp32 rel err {err:.2e}", flush=True)
    assert err < 2e-5
