"""GPU parity of the IF-RK4 path (dns-deform.f90:160-237) and the initial
condition against the oracle, through the C ABI.

Bars: FP64 GPU vs FP64 oracle 1e-11 relative (same algorithm, different FFT
rounding); FP32 GPU vs FP64 oracle on the 64^3 forced-DNS case (100 steps,
dt = 0.002, nu = 0.01): E(k) within 1e-5 relative on every shell holding at
least 1e-6 of the peak shell energy, energy and dissipation within 1e-5,
max |div u| at FP32 round-off (< 1e-4 of max |w|)."""
import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("prec,rt", [(scl.DP, 1e-13), (scl.SP, 2e-6)])
def test_initialize_matches_oracle(prec, rt):
    pr = "sp" if prec == scl.SP else "dp"
    for seed in (1, -2):
        got = scl.initialize(32, 32, 32, seed, 4.0, 1.0, prec)
        want = orc.initialize(32, 32, 32, seed, 4.0, 1.0, prec="dp")
        for a, b in zip(got, want):
            assert relerr(a, b) < rt


@pytest.mark.parametrize("forced", [0, 1])
def test_rk4step_fp64_matches_oracle(forced):
    n = 32
    print(f"case rk4step dp forced={forced}", flush=True)
    p = scl.Plan3D(n, n, n, scl.DP)
    v = [a.copy() for a in orc.initialize(n, n, n, 1, 4.0, 1.0)]
    d = orc.DNS(orc.DNSParams(nx=n, nu=0.01, dt=0.002, forced=bool(forced), fkmax=2.0, eps=0.1))
    w = tuple(v)
    for _ in range(3):
        w = d.rk4_step(*w)
        p.rk4step(*v, 0.01, 0.002, forced, 2.0, 0.1)
    for a, b in zip(v, w):
        assert relerr(a, b) < 1e-11


@pytest.mark.parametrize("use_graph", [False, True])
def test_device_solver_fp64(use_graph):
    n = 32
    p = scl.Plan3D(n, n, n, scl.DP)
    s = scl.DNS(p, nu=0.01, dt=0.002, forced=1, fkmax=2.0, eps=0.1, use_graph=use_graph)
    v0 = orc.initialize(n, n, n, 1, 4.0, 1.0)
    s.set_velocity(*[a.copy() for a in v0])
    s.step(5)
    s.sync()
    d = orc.DNS(orc.DNSParams(nx=n))
    w = v0
    for _ in range(5):
        w = d.rk4_step(*w)
    for a, b in zip(s.get_velocity(), w):
        assert relerr(a, b) < 1e-11
    assert abs(s.time - 5 * 0.002) < 1e-15


def test_graph_replay_is_bitwise_equal_to_eager():
    n = 32
    p = scl.Plan3D(n, n, n, scl.SP)
    out = []
    for g in (False, True):
        s = scl.DNS(p, nu=0.01, dt=0.002, forced=1, use_graph=g)
        s.initialize(1, 4.0, 1.0)
        for _ in range(3):  # replays with host work in between
            s.step(2)
            s.sync()
            _ = s.spectrum()
        out.append(s.get_velocity())
        s.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n", [32, 64])
def test_steps_do_not_depend_on_grouping(n):
    """The stage chain carries the z-inverse of v from one step to the next
    (stage 4's fused z-pass): 6 steps in one call, in calls of 2, 1 and 3
    (graph and eager, with diagnostics and get_velocity in between) and in a
    timed run must give the same bits; setting the state restarts the chain
    from a z-inverse pass."""
    print(f"case step grouping n={n}", flush=True)
    p = scl.Plan3D(n, n, n, scl.SP)
    v0 = scl.initialize(n, n, n, 1, 4.0, 1.0, scl.SP)
    runs = []
    for g, groups in ((True, [6]), (True, [2, 1, 3]), (False, [6]), (False, [1, 1, 4]), (None, [3, 3])):
        s = scl.DNS(p, nu=0.01, dt=0.002, forced=1, use_graph=bool(g))
        s.set_velocity(*[a.copy() for a in v0])
        for k in groups:
            if g is None:
                s.step_timed(k)
            else:
                s.step(k)
            s.sync()
            _ = s.spectrum(), s.diagnostics(), s.get_velocity()
        runs.append(s.get_velocity())
        # a new state restarts the chain: the same 6 steps from v0 again
        s.set_velocity(*[a.copy() for a in v0])
        s.step(6)
        s.sync()
        for a, b in zip(runs[-1], s.get_velocity()):
            np.testing.assert_array_equal(a, b)
        s.close()
    for r in runs[1:]:
        for a, b in zip(runs[0], r):
            np.testing.assert_array_equal(a, b)
    p.close()


def test_helical_forcing_solver_fp64():
    n = 32
    nu, dt, eps, eta = 0.01, 0.002, 0.1, 0.05
    p = scl.Plan3D(n, n, n, scl.DP)
    s = scl.DNS(p, nu=nu, dt=dt, forced=2, fkmax=2.0, eps=eps, eta=eta)
    v0 = orc.initialize(n, n, n, 1, 4.0, 1.0)
    s.set_velocity(*[a.copy() for a in v0])
    s.step(3)
    # oracle: the dns-deform step with force_heli in the RHS
    d = orc.DNS(orc.DNSParams(nx=n, nu=nu, dt=dt, forced=False))
    base_rhs = d.rhs

    def rhs(vx, vy, vz):
        om = orc.vorticity(vx, vy, vz, d.kx, d.ky, d.kz)
        nl = orc.convec_dns(vx, vy, vz, *om, n, n, n)
        f = orc.force_heli(vx, vy, vz, *om, d.k2, 2.0, eps, eta)
        w = [a + b for a, b in zip(nl, f)]
        for a in w:
            orc.symmetrize3d(a)
        orc.projection(*w, d.kx, d.ky, d.kz)
        return tuple(w)

    d.rhs = rhs
    w = v0
    for _ in range(3):
        w = d.rk4_step(*w)
    d.rhs = base_rhs
    for a, b in zip(s.get_velocity(), w):
        assert relerr(a, b) < 1e-11


def test_forced_dns_64_fp32_vs_fp64_oracle():
    """The 64^3 forced-DNS case: 100 IF-RK4 steps in FP32 on the GPU vs the
    FP64 oracle from the same (FP64-generated) initial field."""
    n, nsteps = 64, 100
    print("case dns64 fp32 100 steps", flush=True)
    v0 = orc.initialize(n, n, n, 1, 4.0, 1.0)
    p = scl.Plan3D(n, n, n, scl.SP)
    s = scl.DNS(p, nu=0.01, dt=0.002, forced=1, fkmax=2.0, eps=0.1)
    s.set_velocity(*[a.astype(np.complex64) for a in v0])
    s.step(nsteps)
    s.sync()
    d = orc.DNS(orc.DNSParams(nx=n, nu=0.01, dt=0.002, fkmax=2.0, eps=0.1))
    w = v0
    for _ in range(nsteps):
        w = d.rk4_step(*w)
    ek_ref = orc.spectrum(*w, d.k2)
    ek = s.spectrum()
    sel = ek_ref >= 1e-6 * ek_ref.max()
    err = np.abs(ek[sel] - ek_ref[sel]) / ek_ref[sel]
    print("E(k) max rel err", err.max(), "shells", sel.sum(), flush=True)
    assert err.max() < 1e-5
    dg = s.diagnostics()
    e_ref = orc.energy(*w)
    eps_ref = orc.dissipation(*w, d.k2, 0.01)
    print("energy", dg["energy"], e_ref, "diss", dg["dissipation"], eps_ref, "divmax", dg["divmax"], flush=True)
    assert abs(dg["energy"] - e_ref) < 1e-5 * e_ref
    assert abs(dg["dissipation"] - eps_ref) < 1e-5 * eps_ref
    wmax = np.abs(orc.c2r(orc.vorticity(*w, d.kx, d.ky, d.kz)[0], n)).max()
    assert dg["divmax"] < 1e-4 * wmax
    assert dg["nonfinite"] == 0


def test_diagnostics_match_oracle_fp64():
    n = 32
    p = scl.Plan3D(n, n, n, scl.DP)
    s = scl.DNS(p)
    v = orc.initialize(n, n, n, -4, 4.0, 1.0)
    s.set_velocity(*[a.copy() for a in v])
    kx, ky, kz, k2 = orc.wavenumber(n, n, n)
    np.testing.assert_allclose(s.spectrum(), orc.spectrum(*v, k2), rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(s.helicity_spectrum(), orc.helicity_spectrum(*v, kx, ky, kz, k2), rtol=1e-10, atol=1e-14)
    dg = s.diagnostics()
    assert np.isclose(dg["energy"], orc.energy(*v), rtol=1e-13)
    assert np.isclose(dg["dissipation"], orc.dissipation(*v, k2, 0.01), rtol=1e-13)
    assert np.isclose(dg["enstrophy"], orc.enstrophy_total(*v, k2), rtol=1e-13)
    assert dg["divmax"] < 1e-12
    # helicity sum' 2 Re(w . v*) and its dissipation 2 nu sum' k^2 (...): the
    # shell spectrum above sums to the first
    w = orc.vorticity(*v, kx, ky, kz)
    h = 2 * np.real(sum(a * np.conj(b) for a, b in zip(w, v)))
    hw = np.ones(k2.shape)
    hw[:, :, 0] = 0.5
    k2i = k2.astype(np.float64)
    k2i[0, 0, 0] = 0
    assert np.isclose(dg["helicity"], (hw * h).sum(), rtol=1e-12)
    assert np.isclose(dg["helicity_dissipation"], 2 * 0.01 * (hw * k2i * h).sum(), rtol=1e-12)
