"""GPU parity and throughput of the batched 2D vorticity ensembles
(spectral_dns_2d.f90:94,161-188 RHS; IF-RK4 stepper) against oracle/vort2d.py.

Bars: FP64 initial fields within 1e-13 and RHS within 1e-11 of the FP64
oracle (relative to the field's max); FP64 steps within 1e-10 after 5 steps;
FP32 within 1e-5 (one RHS) and 1e-4 (50 steps) of the FP64 oracle; batched,
per-member and eager launches bit-identical in FP32 under one schedule; the
batched launch of 192 x 256^2 + 48 x 512^2 + 16 x 1024^2 faster than
per-member launches of the same members."""
import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import vort2d as v2

pytestmark = pytest.mark.gpu

KW = dict(dt=5e-4, c_hyper=1e-5, n_hyper=8, damp=0.1, fkmin=4.0, fkmax=5.0, eps=0.01)


def relerr(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _oracle(n, **kw):
    return v2.Vort2D(v2.Vort2DParams(n=n, **kw))


def _members(e, groups):
    return [e.get_member(g, j) for g, (_, m) in enumerate(groups) for j in range(m)]


@pytest.mark.parametrize("prec", [scl.DP, scl.SP])
def test_initialize_and_diagnostics(prec):
    groups = [(32, 2), (64, 3)]
    e = scl.Ensemble2D(groups, prec=prec, **KW)
    e.initialize(seed0=-5, k0=8.0, energy=0.5)
    d = e.diagnostics()
    q = 0
    for g, (n, m) in enumerate(groups):
        k2 = v2.wavenumber2d(n)[2]
        for j in range(m):
            want = v2.initialize2d(n, -5 - q, 8.0, 0.5)
            got = e.get_member(g, j)
            err = relerr(got, want)
            assert err < (1e-13 if prec == scl.DP else 1e-7), (n, j, err)
            assert np.isclose(d[q, 0], v2.energy2d(got, k2), rtol=1e-12)
            assert np.isclose(d[q, 1], v2.enstrophy2d(got), rtol=1e-12)
            q += 1
    assert np.allclose(d[:, 0], 0.5, rtol=1e-6)
    e.close()


@pytest.mark.parametrize("n", [32, 64, 256])
def test_rhs_fp64(n):
    print(f"case ens2d rhs fp64 n={n}", flush=True)
    e = scl.Ensemble2D([(n, 3)], prec=scl.DP, **KW)
    w = v2.initialize2d(n, -11, 8.0, 0.5)
    got = e.rhs(0, 1, w.copy())
    want = _oracle(n, **KW).rhs(w)
    err = relerr(got, want)
    print("rhs fp64 rel err", err, flush=True)
    assert err < 1e-11
    # kx = 0 column is Hermitian, the ky = -n/2 row and kx = n/2 column are zero
    assert np.array_equal(got[n // 2 + 1:, 0], np.conj(got[1:n // 2, 0][::-1]))
    assert not np.any(got[n // 2, :]) and not np.any(got[:, n // 2])
    e.close()


@pytest.mark.parametrize("per_member", [False, True])
def test_steps_fp64(per_member):
    groups = [(32, 3), (64, 2)]
    kw = dict(KW, dt=2e-3)
    print(f"case ens2d steps fp64 per_member={per_member}", flush=True)
    e = scl.Ensemble2D(groups, prec=scl.DP, per_member=per_member, **kw)
    e.initialize(seed0=-21)
    e.step(5)
    got = _members(e, groups)
    q = 0
    for g, (n, m) in enumerate(groups):
        o = _oracle(n, **kw)
        for j in range(m):
            w = v2.initialize2d(n, -21 - q, 8.0, 0.5)
            for _ in range(5):
                w = o.rk4_step(w)
            err = relerr(got[q], w)
            assert err < 1e-10, (n, j, err)
            q += 1
    assert np.isclose(e.time, 5 * 2e-3)
    e.close()


def test_fp32_batched_per_member_eager_bitwise(monkeypatch):
    groups = [(64, 3), (128, 2)]
    print("case ens2d fp32 batched/per-member/eager/forked bitwise", flush=True)
    runs = []
    for per_member, use_graph, fork in ((False, True, 0), (False, False, 0), (True, True, 0), (True, False, 0),
                                        (False, True, 1), (False, False, 1)):
        monkeypatch.setenv("SCL_ENS2D_FORK", str(fork))  # a graph branch per size group
        e = scl.Ensemble2D(groups, prec=scl.SP, per_member=per_member, use_graph=use_graph, **KW)
        e.initialize(seed0=-31)
        e.step(4)
        e.step(11)
        runs.append(_members(e, groups))
        e.close()
    for r in runs[1:]:
        for a, b in zip(runs[0], r):
            assert np.array_equal(a, b)


def test_schedules_fp32_rhs():
    n = 256
    w = v2.initialize2d(n, -41, 8.0, 0.5)
    want = _oracle(n, **KW).rhs(w)
    e = scl.Ensemble2D([(n, 2)], prec=scl.SP, **KW)
    for sched in [(8, 8, 0), (8, 16, 0), (16, 8, 1), (16, 16, 0), (8, 8, 1)]:
        print("case ens2d schedule", sched, flush=True)
        e.set_schedule(0, *sched)
        assert e.get_schedule(0) == sched
        got = e.rhs(0, 1, w.astype(np.complex64))
        err = relerr(got, want)
        print("schedule", sched, "fp32 rhs rel err", err, flush=True)
        assert err < 1e-5
    e.close()


def test_fp32_run_vs_fp64_oracle():
    n, m, nsteps = 128, 2, 50
    print("case ens2d fp32 50 steps vs fp64 oracle", flush=True)
    e = scl.Ensemble2D([(n, m)], prec=scl.SP, **KW)
    e.initialize(seed0=-51)
    e.step(nsteps)
    d = e.diagnostics()
    o = _oracle(n, **KW)
    k2 = v2.wavenumber2d(n)[2]
    for j in range(m):
        w = v2.initialize2d(n, -51 - j, 8.0, 0.5)
        for _ in range(nsteps):
            w = o.rk4_step(w)
        got = e.get_member(0, j)
        err = relerr(got, w)
        ee = abs(d[j, 0] - v2.energy2d(w, k2)) / v2.energy2d(w, k2)
        ez = abs(d[j, 1] - v2.enstrophy2d(w)) / v2.enstrophy2d(w)
        print(f"member {j}: field rel err {err:.2e} energy {ee:.2e} enstrophy {ez:.2e}", flush=True)
        assert err < 1e-4 and ee < 1e-5 and ez < 1e-5
    e.close()


def test_batched_faster_than_per_member():
    groups = [(256, 192), (512, 48), (1024, 16)]
    print("case ens2d 192x256^2 + 48x512^2 + 16x1024^2: tune, batched vs per-member", flush=True)
    eb = scl.Ensemble2D(groups, prec=scl.SP, **KW)
    eb.initialize(seed0=-1)
    best = eb.tune(reps=3)
    sched = [eb.get_schedule(g) for g in range(len(groups))]
    print("tuned schedules", sched, "ms per RHS", best, flush=True)
    eb.step(5)
    tb = eb.step_timed(50) / 50
    ep = scl.Ensemble2D(groups, prec=scl.SP, per_member=True, **KW)
    ep.initialize(seed0=-1)
    for g, s in enumerate(sched):
        ep.set_schedule(g, *s)
    ep.step(2)
    tp = ep.step_timed(10) / 10
    print(f"ms/step batched {tb:.3f} per-member {tp:.3f} speedup {tp / tb:.1f}", flush=True)
    d = eb.diagnostics()
    assert np.all(np.isfinite(d))
    assert tb < tp
    eb.close()
    ep.close()
