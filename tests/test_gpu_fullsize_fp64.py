"""The 512^3 forced DNS that bench.py times (dns-deform.f90's IF-RK4, nu =
0.01, dt = 0.002, force() with fkmax = 2 and eps = 0.1, from initialize(1,
akp = 4, u0 = 1)) in FP32 against the same run on an FP64 plan, both on the
GPU, after 20 steps (graph replay).  The FP64 plan's 512-point kernels are
held to the FP64 oracle at 1e-11 on thin grids (test_gpu_zlines,
test_gpu_wide_rows), so this compares the FP32 kernels at their full size
with a reference of known accuracy.

Bars: the velocity fields within 1e-5 of their largest entry, the energy,
dissipation and enstrophy within 1e-5 relative, E(k) within 1e-4 relative
on every shell holding more than 1e-12 of the total (FP32 rounding, 6e-8,
over 20 steps of a smooth, well-resolved flow: about 1e-6 is expected)."""
import numpy as np
import pytest

import speccodelib_amd as scl

pytestmark = pytest.mark.gpu

N, STEPS = 512, 20


def _run(prec):
    p = scl.Plan3D(N, N, N, prec)
    s = scl.DNS(p, nu=0.01, dt=0.002, forced=1, fkmax=2.0, eps=0.1, use_graph=True)
    s.initialize(1, 4.0, 1.0)
    s.step(STEPS)
    s.sync()
    out = dict(ek=s.spectrum(), diag=s.diagnostics(), v=[np.asarray(a, np.complex128) for a in s.get_velocity()])
    s.close()
    p.close()
    return out


def test_512_fp32_against_fp64_plan():
    print(f"case {N}^3 forced DNS, {STEPS} steps, FP32 vs FP64 plan", flush=True)
    a = _run(scl.SP)
    print("FP32 run done", flush=True)
    b = _run(scl.DP)
    print("FP64 run done", flush=True)
    ferr = max(np.abs(x - y).max() / np.abs(y).max() for x, y in zip(a["v"], b["v"]))
    print(f"velocity rel err {ferr:.2e}", flush=True)
    assert ferr < 1e-5
    for key in ("energy", "dissipation", "enstrophy"):
        r = abs(a["diag"][key] - b["diag"][key]) / abs(b["diag"][key])
        print(f"{key}: {a['diag'][key]:.9e} vs {b['diag'][key]:.9e}, rel {r:.2e}", flush=True)
        assert r < 1e-5
    ek32, ek64 = a["ek"], b["ek"]
    sel = ek64 > 1e-12 * ek64.sum()
    r = np.abs(ek32[sel] - ek64[sel]) / ek64[sel]
    print(f"E(k): {sel.sum()} shells, max rel {r.max():.2e} (shell {np.flatnonzero(sel)[r.argmax()] + 1})", flush=True)
    assert r.max() < 1e-4
