"""1024^3 FP32 forced DNS with helical forcing (force_heli.f90:1-57) on one
GPU: seeded k^4 exp(-2 (k/4)^2) initial field (u0 = 1, iseed = 1),
nu = 2.5e-4, dt = 2.5e-4, forcing on k^2 < fkmax^2 = 4 with energy input
eps = 0.1 and helicity input eta = 2 k_f eps (relative helicity 1 at
k_f = 2), 2/3 truncation, 50 IF-RK4 steps.

No oracle runs at this size, so the pins are the budgets the discrete
system obeys up to time-integration error:
  E(T) - E(0) = eps T - int 2 nu sum' k^2 |v|^2 dt,
  H(T) - H(0) = eta T - int 2 nu sum' k^2 h dt,
with the integrals by the trapezoid rule over 5-step intervals, both within
1% of the injected amount; the field stays finite and divergence-free, and
E(k), H(k) come from the FP64 shell reductions."""
import time

import numpy as np
import pytest

import speccodelib_amd as scl

pytestmark = pytest.mark.gpu


def test_helical_forced_dns_1024():
    n, nsteps, dt, nu, eps, kf = 1024, 50, 2.5e-4, 2.5e-4, 0.1, 2.0
    eta = 2 * kf * eps
    print("case helical dns 1024^3 fp32, 50 steps", flush=True)
    p = scl.Plan3D(n, n, n, scl.SP)
    d = scl.DNS(p, nu=nu, dt=dt, forced=2, fkmax=kf, eps=eps, eta=eta)
    t0 = time.perf_counter()
    d.initialize(1, 4.0, 1.0)
    d.sync()
    print(f"initial field {time.perf_counter() - t0:.1f} s", flush=True)
    hist = [d.diagnostics()]
    every = 5
    step_s = 0.0
    for k in range(nsteps // every):
        t0 = time.perf_counter()
        d.step(every)
        d.sync()
        step_s += time.perf_counter() - t0
        hist.append(d.diagnostics())
        print(f"step {(k + 1) * every}: E {hist[-1]['energy']:.8f} H {hist[-1]['helicity']:.8f}", flush=True)
    T = nsteps * dt
    E = np.array([h["energy"] for h in hist])
    D = np.array([h["dissipation"] for h in hist])
    H = np.array([h["helicity"] for h in hist])
    HD = np.array([h["helicity_dissipation"] for h in hist])
    tau = every * dt
    intD = tau * (D.sum() - 0.5 * (D[0] + D[-1]))
    intHD = tau * (HD.sum() - 0.5 * (HD[0] + HD[-1]))
    e_err = abs((E[-1] - E[0]) - (eps * T - intD)) / (eps * T)
    h_err = abs((H[-1] - H[0]) - (eta * T - intHD)) / (eta * T)
    F = 8 * (n // 2 + 1) * n * n
    ms = step_s * 1e3 / nsteps
    frac = 198 * F / (ms * 1e-3) / 8e12
    ek, hk = d.spectrum(), d.helicity_spectrum()
    print(f"1024^3: {ms:.1f} ms/step ({1e3 / ms:.2f} steps/s, {frac:.1%} of HBM roofline at 198 F/step); "
          f"energy budget err {e_err:.2e}, helicity budget err {h_err:.2e}; divmax {hist[-1]['divmax']:.2e}", flush=True)
    print("E(k) k=1..8:", np.array2string(ek[:8], precision=4), flush=True)
    print("H(k) k=1..8:", np.array2string(hk[:8], precision=4), flush=True)
    assert all(h["nonfinite"] == 0 for h in hist)
    assert np.all(np.isfinite(ek)) and np.all(np.isfinite(hk))
    assert hist[-1]["divmax"] < 1e-3
    assert e_err < 1e-2 and h_err < 1e-2
    d.close()
    p.close()
