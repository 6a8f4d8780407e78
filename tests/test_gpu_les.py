"""GPU parity of the LES (dynsmag.f90, smag.f90) and passive-scalar
(convecphi_dns.f90) paths against the oracle, through the C ABI.

Bars: FP64 plans vs the FP64 oracle at 1e-11 relative (routines) and 1e-10
(solver steps); the 256^3 static-Smagorinsky LES + scalar case (C_s = 0.17,
Delta = 2 pi/256, nu = 1e-4, Sc = 0.7, unit mean gradient, 500 steps of
dt = 0.001) in FP32 vs the same run in FP64 on the GPU: E(k) within 1e-4 on
shells holding >= 1e-6 of the peak, energy, dissipation and scalar variance
within 1e-4 (errors printed)."""
import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import les_scalar as les
from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _k3(n):
    kx, ky, kz, k2 = orc.wavenumber(n, n, n)
    shape = k2.shape
    return (np.ascontiguousarray(np.broadcast_to(kx[None, None, :], shape)),
            np.ascontiguousarray(np.broadcast_to(ky[None, :, None], shape)),
            np.ascontiguousarray(np.broadcast_to(kz[:, None, None], shape)))


def test_smag_and_static_dynsmag_fp64():
    n = 32
    p = scl.Plan3D(n, n, n, scl.DP)
    kx, ky, kz, k2 = orc.wavenumber(n, n, n)
    v = orc.initialize(n, n, n, -7, 4.0, 1.0)
    delta = 2 * np.pi / n
    got = p.smag(*v, kx, ky, kz, delta)
    want = les.smag_tau(*v, kx, ky, kz, float(np.float32(0.026)), delta, (n, n, n))
    for a, b in zip(got, want):
        assert relerr(a, b) < 1e-11
    g = les.g_test(k2, 2 * (n // 2) / 3 / 2)
    tau, mm, lm, c = p.dynsmag(*v, *_k3(n), g, delta, 2 * delta, cs2=0.0289, update=False)
    want = les.smag_tau(*v, kx, ky, kz, 0.0289, delta, (n, n, n))
    for a, b in zip(tau, want):
        assert relerr(a, b) < 1e-11


def test_dynsmag_update_fp64():
    n = 32
    p = scl.Plan3D(n, n, n, scl.DP)
    kx, ky, kz, k2 = orc.wavenumber(n, n, n)
    v = orc.initialize(n, n, n, -8, 4.0, 1.0)
    delta = np.pi / (2 * (n // 2) / 3)
    g = les.g_test(k2, 2 * (n // 2) / 3 / 2)
    tau, mm, lm, c = p.dynsmag(*v, *_k3(n), g, delta, 2 * delta, update=True)
    wt, wmm, wlm, wc = les.dynsmag_update(*v, kx, ky, kz, g, delta, 2 * delta, (n, n, n))
    print("cs2", c, wc, "mm", mm, wmm, "lm", lm, wlm, flush=True)
    assert np.isclose(mm, wmm, rtol=1e-11) and np.isclose(lm, wlm, rtol=1e-10) and np.isclose(c, wc, rtol=1e-10)
    for a, b in zip(tau, wt):
        assert relerr(a, b) < 1e-10


@pytest.mark.parametrize("prec", [scl.SP, scl.DP])
def test_convecphi_dns(prec):
    n = 32
    pr = "sp" if prec == scl.SP else "dp"
    p = scl.Plan3D(n, n, n, prec)
    kx, ky, kz, k2 = orc.wavenumber(n, n, n, pr)
    v = orc.initialize(n, n, n, -9, 4.0, 1.0, prec=pr)
    phi = orc.initialize(n, n, n, -10, 3.0, 1.0, prec=pr)[0]
    got = p.convecphi_dns(*v, phi, kx, ky, kz)
    want = les.convecphi_dns(*[a.astype(np.complex128) for a in v], phi.astype(np.complex128),
                             kx.astype(np.float64), ky.astype(np.float64), kz.astype(np.float64), (n, n, n))
    assert relerr(got, want) < (1e-11 if prec == scl.DP else 3e-6)


def test_les_scalar_rhs_fp64():
    """scl_dns_rhs of the LES + scalar solver (the RHS mode of its fused
    z-passes) against the oracle's N - nu k^2 y (scalar: nu/Sc), 1e-11."""
    n, nu, sc = 32, 1e-3, 0.7
    delta = 2 * np.pi / n
    print("case les+scalar rhs fp64", flush=True)
    prm = les.LESParams(nx=n, nu=nu, dt=0.002, forced=True, fkmax=2.5, eps=0.1, cs2=0.0289, delta=delta,
                        scalar=True, G=1.0, Sc=sc)
    o = les.LESScalar(prm)
    p = scl.Plan3D(n, n, n, scl.DP)
    s = scl.DNS(p, nu=nu, dt=0.002, forced=1, fkmax=2.5, eps=0.1, sgs=1, cs2=0.0289, delta=delta, scalar=True,
                grad=1.0, schmidt=sc)
    v = orc.initialize(n, n, n, 1, 4.0, 1.0)
    phi = orc.initialize(n, n, n, -10, 3.0, 1.0)[0]
    got = s.rhs(*[a.copy() for a in v], phi.copy())
    want = o.rhs(*[a.copy() for a in v], phi.copy())
    k2 = orc.wavenumber(n, n, n)[3].astype(np.float64)
    k2[0, 0, 0] = 0
    lin = [nu * k2] * 3 + [nu / sc * k2]
    for a, b, y, c in zip(got, want, (*v, phi), lin):
        assert relerr(a, b - c * y) < 1e-11
    s.close()
    p.close()


def _les_pair(n, dynamic, nsteps, iupdate=2):
    nu, dt = 1e-3, 0.002
    delta = 2 * np.pi / n
    kc = 2 * (n // 2) / 3
    prm = les.LESParams(nx=n, nu=nu, dt=dt, forced=True, fkmax=2.5, eps=0.1, cs2=0.0289, delta=delta,
                        dynamic=dynamic, iupdate=iupdate, scalar=True, G=1.0, Sc=0.7)
    o = les.LESScalar(prm)
    p = scl.Plan3D(n, n, n, scl.DP)
    s = scl.DNS(p, nu=nu, dt=dt, forced=1, fkmax=2.5, eps=0.1, sgs=2 if dynamic else 1, cs2=0.0289, delta=delta,
                delta_test=2 * delta, k_test=kc / 2, iupdate=iupdate, scalar=True, grad=1.0, schmidt=0.7)
    v = orc.initialize(n, n, n, 1, 4.0, 1.0)
    phi = np.zeros_like(v[0])
    s.set_velocity(*[a.copy() for a in v])
    s.set_scalar(phi.copy())
    s.step(nsteps)
    w = (*v, phi)
    for _ in range(nsteps):
        w = o.rk4_step(*w)
    return s, w, o


@pytest.mark.parametrize("dynamic", [False, True])
def test_les_scalar_solver_fp64(dynamic):
    n, nsteps = 32, 3
    print(f"case les solver dynamic={dynamic}", flush=True)
    s, w, o = _les_pair(n, dynamic, nsteps)
    for a, b in zip(s.get_velocity(), w[:3]):
        assert relerr(a, b) < 1e-10
    assert relerr(s.get_scalar(), w[3]) < 1e-10
    if dynamic:
        assert np.isclose(s.diagnostics()["cs2"], o.cs2, rtol=1e-9)


def test_les_scalar_256_fp32_vs_fp64():
    """Static Smagorinsky LES + passive scalar at 256^3, FP32 vs FP64 on the GPU."""
    n, nsteps = 256, 500
    print("case les256 fp32 vs fp64, 500 steps", flush=True)
    kw = dict(nu=1e-4, dt=0.001, forced=1, fkmax=2.5, eps=0.1, sgs=1, cs2=0.17**2, delta=2 * np.pi / n, scalar=True,
              grad=1.0, schmidt=0.7)
    v = orc.initialize(n, n, n, 1, 4.0, 1.0)
    phi = np.zeros_like(v[0])
    res = {}
    for prec in (scl.DP, scl.SP):
        p = scl.Plan3D(n, n, n, prec)
        s = scl.DNS(p, **kw)
        cd = scl._cdtype(prec)
        s.set_velocity(*[a.astype(cd) for a in v])
        s.set_scalar(phi.astype(cd))
        s.step(nsteps)
        s.sync()
        res[prec] = (s.spectrum(), s.diagnostics())
        s.close()
        p.close()
    ek64, d64 = res[scl.DP]
    ek32, d32 = res[scl.SP]
    sel = ek64 >= 1e-6 * ek64.max()
    e_ek = (np.abs(ek32[sel] - ek64[sel]) / ek64[sel]).max()
    errs = {k: abs(d32[k] - d64[k]) / abs(d64[k]) for k in ("energy", "dissipation", "scalar_variance", "scalar_dissipation")}
    print("LES256 FP32 vs FP64: E(k) max rel err", e_ek, {k: f"{v:.2e}" for k, v in errs.items()},
          "energy", d64["energy"], "scalar var", d64["scalar_variance"], flush=True)
    assert d32["nonfinite"] == 0 and d64["nonfinite"] == 0
    assert e_ek < 1e-4
    assert all(v < 1e-4 for v in errs.values())
