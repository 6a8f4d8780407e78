"""GPU checks of the RHS-callback form (rk5.f90:1-12 derivs convention):
scl_dns_rhs against the FP64 oracle's N(v) - nu k^2 v (1e-11 relative), the
Fortran-callable scl_derivs_ against scl_dns_rhs (bit-identical), and the
device IF-RK4 driver with the built-in spectral callback against the
solver's own steps (1e-12 relative, FP64; RHS truncation on in both, so the
two are the same scheme)."""
import ctypes

import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_dns_rhs_matches_oracle():
    n, nu = 32, 0.01
    print("case derivs scl_dns_rhs fp64", flush=True)
    p = scl.Plan3D(n, n, n, scl.DP)
    d = scl.DNS(p, nu=nu, dt=0.002, forced=1, fkmax=2.0, eps=0.1)
    v = orc.initialize(n, n, n, -4, 4.0, 1.0)
    got = d.rhs(*[a.copy() for a in v])
    o = orc.DNS(orc.DNSParams(nx=n, nu=nu, dt=0.002, forced=True, fkmax=2.0, eps=0.1))
    N = o.rhs(*v)
    k2 = orc.wavenumber(n, n, n)[3].astype(np.float64)
    k2[0, 0, 0] = 0
    for a, b, c in zip(got, N, v):
        assert relerr(a, b - nu * k2 * c) < 1e-11
    # the Fortran-callable form over the concatenated real(dp) state
    lib = scl.lib()
    scl._check(lib.scl_derivs_bind(d.h))
    y = np.concatenate([a.view(np.float64).ravel() for a in v])
    dydx = np.zeros_like(y)
    x = ctypes.c_double(0.0)
    ny = ctypes.c_int32(y.size)
    lib.scl_derivs_(ctypes.byref(x), y.ctypes.data, dydx.ctypes.data, ctypes.byref(ny))
    want = np.concatenate([a.view(np.float64).ravel() for a in got])
    assert np.array_equal(dydx, want)
    scl._check(lib.scl_derivs_bind(None))
    d.close()
    p.close()


def test_ifrk4_driver_with_builtin_callback():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible to torch")
    n, nsteps = 32, 5
    print("case derivs ifrk4 driver vs solver steps", flush=True)
    kw = dict(nu=0.01, dt=0.002, forced=1, fkmax=2.0, eps=0.1, dealias_rhs=True, use_graph=False)
    p = scl.Plan3D(n, n, n, scl.DP)
    v = orc.initialize(n, n, n, -6, 4.0, 1.0)
    d1 = scl.DNS(p, **kw)
    d1.set_velocity(*[a.copy() for a in v])
    d1.step(nsteps)
    ref = d1.get_velocity()
    d2 = scl.DNS(p, **kw)
    d2.set_velocity(*[a.copy() for a in v])
    ptr, ne = d2.state()
    rate = torch.empty(ne, dtype=torch.float64, device="cuda")
    d2.rates(rate.data_ptr())
    torch.cuda.synchronize()
    scl.ifrk4_integrate(scl.DP, ne, ptr, rate.data_ptr(), 0.0, 0.002, nsteps, scl.dns_derivs_cb_ptr(), d2.h.value)
    got = d2.get_velocity()
    err = max(relerr(a, b) for a, b in zip(got, ref))
    print("driver vs solver rel err", err, flush=True)
    assert err < 1e-12
    d1.close()
    d2.close()
    p.close()
