"""The two DNS runs bench.py times, in FP32 against the same run on an FP64
plan, both on the GPU (graph replay): the 512^3 forced DNS (dns-deform.f90's
IF-RK4, nu = 0.01, dt = 0.002, force() with fkmax = 2 and eps = 0.1, from
initialize(1, akp = 4, u0 = 1)) after 20 steps, and the 1024^3 helical-forced
DNS (nu = dt = 2.5e-4, force_heli with fkmax = 2, eps = 0.1, eta = 0.4, the
same start) after 10 steps.  The FP64 plan's 512- and 1024-point kernels are
held to the FP64 oracle at 1e-11 on thin grids (test_gpu_zlines,
test_gpu_wide_rows), so this compares the FP32 kernels at their full size
with a reference of known accuracy.

Bars: the velocity fields within 1e-5 of their largest entry, the energy,
dissipation and enstrophy (and at 1024^3 the helicity, on the scale
max(|H|, sqrt(E Omega))) within 1e-5 relative, E(k) within 1e-4 relative on every shell holding more than 1e-12
of the total (FP32 rounding, 6e-8, over tens of steps of a smooth,
well-resolved flow: about 1e-6 is expected).

The 2D ensembles bench.py times (192 x 256^2 + 48 x 512^2 + 16 x 1024^2
members, spectral_dns_2d.f90's RHS with IF-RK4, the same seeded start)
likewise after 50 batched steps: every member's energy and enstrophy
within 1e-5 relative, the first and last member of each size within 1e-4
of their largest entry."""
import numpy as np
import pytest

import speccodelib_amd as scl

pytestmark = pytest.mark.gpu

FORCED = dict(nu=0.01, dt=0.002, forced=1, fkmax=2.0, eps=0.1)
HELICAL = dict(nu=2.5e-4, dt=2.5e-4, forced=2, fkmax=2.0, eps=0.1, eta=0.4)


def _run(n, prec, steps, kw):
    p = scl.Plan3D(n, n, n, prec)
    s = scl.DNS(p, use_graph=True, **kw)
    s.initialize(1, 4.0, 1.0)
    s.step(steps)
    s.sync()
    out = dict(ek=s.spectrum(), diag=s.diagnostics(), v=s.get_velocity())
    s.close()
    p.close()
    return out


def _compare(n, steps, kw, keys):
    print(f"case {n}^3 DNS {kw}, {steps} steps, FP32 vs FP64 plan", flush=True)
    a = _run(n, scl.SP, steps, kw)
    print("FP32 run done", flush=True)
    b = _run(n, scl.DP, steps, kw)
    print("FP64 run done", flush=True)
    ferr = 0.0
    for x, y in zip(a["v"], b["v"]):
        ferr = max(ferr, np.abs(x.astype(np.complex128) - y).max() / np.abs(y).max())
    print(f"velocity rel err {ferr:.2e}", flush=True)
    assert ferr < 1e-5
    # helicity is measured on its natural scale sqrt(E Omega) (it may be small)
    scale = np.sqrt(b["diag"]["energy"] * b["diag"]["enstrophy"])
    for key in keys:
        den = max(abs(b["diag"][key]), scale) if key == "helicity" else abs(b["diag"][key])
        r = abs(a["diag"][key] - b["diag"][key]) / den
        print(f"{key}: {a['diag'][key]:.9e} vs {b['diag'][key]:.9e}, rel {r:.2e}", flush=True)
        assert r < 1e-5
    ek32, ek64 = a["ek"], b["ek"]
    sel = ek64 > 1e-12 * ek64.sum()
    r = np.abs(ek32[sel] - ek64[sel]) / ek64[sel]
    print(f"E(k): {sel.sum()} shells, max rel {r.max():.2e} (shell {np.flatnonzero(sel)[r.argmax()] + 1})", flush=True)
    assert r.max() < 1e-4


def test_512_fp32_against_fp64_plan():
    _compare(512, 20, FORCED, ("energy", "dissipation", "enstrophy"))


def test_1024_helical_fp32_against_fp64_plan():
    _compare(1024, 10, HELICAL, ("energy", "dissipation", "enstrophy", "helicity"))


ENS_GROUPS = [(256, 192), (512, 48), (1024, 16)]
ENS_KW = dict(dt=5e-4, c_hyper=1e-5, n_hyper=8, damp=0.1, fkmin=4.0, fkmax=5.0, eps=0.01)


def test_ensemble_2d_fp32_against_fp64():
    print("case 2D ensembles 192x256^2 + 48x512^2 + 16x1024^2, 50 steps, FP32 vs FP64", flush=True)
    out = {}
    for prec in (scl.SP, scl.DP):
        e = scl.Ensemble2D(ENS_GROUPS, prec=prec, **ENS_KW)
        e.initialize(seed0=-1, k0=8.0, energy=0.5)
        e.step(50)
        e.sync()
        fields = {(g, j): e.get_member(g, j) for g, (_, m) in enumerate(ENS_GROUPS) for j in (0, m - 1)}
        out[prec] = (e.diagnostics(), fields)
        e.close()
        print(f"prec {prec} done", flush=True)
    (d32, f32), (d64, f64) = out[scl.SP], out[scl.DP]
    r = np.abs(d32 - d64) / np.abs(d64)
    print(f"member energy and enstrophy: max rel {r.max():.2e}", flush=True)
    assert r.max() < 1e-5
    for key, b in f64.items():
        err = np.abs(f32[key].astype(np.complex128) - b).max() / np.abs(b).max()
        print(f"group {key[0]} member {key[1]}: rel err {err:.2e}", flush=True)
        assert err < 1e-4
