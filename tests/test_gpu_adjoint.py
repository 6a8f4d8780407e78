"""GPU checks of the discrete adjoint of the IF-RK4 map (unforced DNS and
static Smagorinsky LES).

Pins: (1) the RHS vector-Jacobian product against central differences of
the FP64 oracle RHS (oracle/les_scalar.py, oracle/spectral3d.py): the
advective part is quadratic, so the difference is exact up to rounding, and
the Smagorinsky part is smooth away from |S| = 0; bar 1e-9 relative.
(2) Gradients of L = sum' h (|u(T) - u_tgt|^2 + k^2 |u(T)|^2) (h = 1/2 on
kx = 0: one half of the mean-square mismatch plus the enstrophy) with
respect to u(0) and cs2 against central differences (eps = 1e-3, FP64) of
the library's own FP64 forward map along seeded random directions; bar 1e-4
relative.  (3) FP32 gradients against the FP64 adjoint.  (4) The torch
autograd wrapper against the C ABI.  (5) The 128^3, 200-step, checkpoint-20
case with C_s = 0.15, with the backward/forward time ratio printed."""
import numpy as np
import pytest

import speccodelib_amd as scl
from oracle import les_scalar as les
from oracle import spectral3d as orc

pytestmark = pytest.mark.gpu


def rdot(a, b):
    return float(sum(np.real(np.conj(x.astype(np.complex128)) * y.astype(np.complex128)).sum() for x, y in zip(a, b)))


def _k2int(n):
    k2 = orc.wavenumber(n, n, n)[3].astype(np.float64)
    k2[0, 0, 0] = 0.0
    return k2


def loss_grad(uT, tgt, k2):
    h = np.ones(k2.shape)
    h[:, :, 0] = 0.5
    L = 0.0
    g = []
    for u, t in zip(uT, tgt):
        u = u.astype(np.complex128)
        t = t.astype(np.complex128)
        L += float((h * (np.abs(u - t) ** 2 + k2 * np.abs(u) ** 2)).sum())
        g.append(2 * h * (u - t) + 2 * h * k2 * u)
    return L, g


def _random_field(n, seed):
    rng = np.random.default_rng(seed)
    shape = (n, n, n // 2 + 1)
    f = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    for a in f:
        orc.symmetrize3d(a)
    return f


@pytest.mark.parametrize("sgs", [0, 1])
def test_vjp_dot_product_fp64(sgs):
    n, cs2 = 32, 0.0225
    print(f"case adjoint vjp fp64 sgs={sgs}", flush=True)
    p = scl.Plan3D(n, n, n, scl.DP)
    a = scl.Adjoint(p, 1, 1, nu=0.005, dt=0.002, sgs=sgs)
    s = orc.initialize(n, n, n, -3, 3.0, 1.0)
    nb = _random_field(n, 5)
    d = _random_field(n, 6)
    # a white-noise direction scaled to the state's norm keeps the central
    # difference of the |S| nonlinearity in its converged range
    sc = np.sqrt(rdot(s, s) / rdot(d, d))
    d = [x * sc for x in d]
    out, dc = a.vjp(s, nb, cs2)
    prm = les.LESParams(nx=n, nu=0.005, dt=0.002, forced=False, cs2=cs2 if sgs else 0.0, delta=2 * np.pi / n,
                        scalar=False)
    o = les.LESScalar(prm)
    phi = np.zeros_like(s[0])
    eps = 1e-6
    rp = o.rhs(*[x + eps * y for x, y in zip(s, d)], phi)[:3]
    rm = o.rhs(*[x - eps * y for x, y in zip(s, d)], phi)[:3]
    jd = [(x - y) / (2 * eps) for x, y in zip(rp, rm)]
    lhs, rhs = rdot(out, d), rdot(nb, jd)
    print("vjp dot", lhs, rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-9 * abs(rhs)
    if sgs:
        o.cs2 = cs2 * (1 + 1e-3)
        rp = o.rhs(*s, phi)[:3]
        o.cs2 = cs2 * (1 - 1e-3)
        rm = o.rhs(*s, phi)[:3]
        dn = [(x - y) / (2e-3 * cs2) for x, y in zip(rp, rm)]
        want = rdot(nb, dn)
        print("dN/dcs2", dc, want, flush=True)
        assert abs(dc - want) <= 1e-9 * abs(want)
    a.close()
    p.close()


def _run_case(n, nsteps, ckpt, sgs, cs2, prec, nu=0.005, dt=0.002, k0=3.0):
    p = scl.Plan3D(n, n, n, prec)
    a = scl.Adjoint(p, nsteps, ckpt, nu=nu, dt=dt, sgs=sgs)
    cd = scl._cdtype(prec)
    u0 = [x.astype(cd) for x in scl.initialize(n, n, n, -1, k0, 1.0, scl.DP)]
    v2 = [x.astype(cd) for x in scl.initialize(n, n, n, -2, k0, 1.0, scl.DP)]
    tgt = a.forward(v2, cs2)
    return p, a, u0, tgt


def _fd_check(n, nsteps, ckpt, sgs, cs2, ndir, tol):
    k2 = _k2int(n)
    p, a, u0, tgt = _run_case(n, nsteps, ckpt, sgs, cs2, scl.DP)
    uT = a.forward(u0, cs2)
    L0, gT = loss_grad(uT, tgt, k2)
    g0, dcs2 = a.backward(gT)
    f, b = a.timing()
    print(f"n={n} steps={nsteps} ckpt={ckpt} sgs={sgs}: L={L0:.6e} forward {f:.1f} ms backward {b:.1f} ms "
          f"ratio {b / f:.2f}", flush=True)
    eps = 1e-3
    worst = 0.0
    for j in range(ndir):
        d = scl.initialize(n, n, n, -(100 + j), 3.0, 1.0, scl.DP)
        Lp = loss_grad(a.forward([x + eps * y for x, y in zip(u0, d)], cs2), tgt, k2)[0]
        Lm = loss_grad(a.forward([x - eps * y for x, y in zip(u0, d)], cs2), tgt, k2)[0]
        fd = (Lp - Lm) / (2 * eps)
        ad = rdot(g0, d)
        err = abs(ad - fd) / abs(fd)
        worst = max(worst, err)
        print(f"  direction {j}: adjoint {ad:.10e} fd {fd:.10e} rel err {err:.2e}", flush=True)
    if sgs:
        ec = 1e-3 * cs2
        Lp = loss_grad(a.forward(u0, cs2 + ec), tgt, k2)[0]
        Lm = loss_grad(a.forward(u0, cs2 - ec), tgt, k2)[0]
        fd = (Lp - Lm) / (2 * ec)
        err = abs(dcs2 - fd) / abs(fd)
        print(f"  dL/dcs2: adjoint {dcs2:.10e} fd {fd:.10e} rel err {err:.2e}", flush=True)
        worst = max(worst, err)
    assert worst < tol
    return p, a, u0, tgt, g0, dcs2


@pytest.mark.parametrize("sgs", [0, 1])
def test_gradient_vs_fd_fp64_small(sgs):
    print(f"case adjoint fd 32^3 sgs={sgs}", flush=True)
    p, a, *_ = _fd_check(32, 10, 4, sgs, 0.0225, 3, 1e-4)
    a.close()
    p.close()


def test_fp32_gradient_vs_fp64():
    n, nsteps, ckpt, cs2 = 32, 10, 4, 0.0225
    print("case adjoint fp32 vs fp64 32^3", flush=True)
    k2 = _k2int(n)
    res = {}
    for prec in (scl.DP, scl.SP):
        p, a, u0, tgt = _run_case(n, nsteps, ckpt, 1, cs2, prec)
        uT = a.forward(u0, cs2)
        _, gT = loss_grad(uT, tgt, k2)
        cd = scl._cdtype(prec)
        res[prec] = a.backward([x.astype(cd) for x in gT])
        a.close()
        p.close()
    g64, c64 = res[scl.DP]
    g32, c32 = res[scl.SP]
    err = max(np.abs(x - y).max() for x, y in zip(g32, g64)) / max(np.abs(y).max() for y in g64)
    print("fp32 vs fp64 gradient rel err", err, "dcs2", c32, c64, flush=True)
    assert err < 1e-4 and abs(c32 - c64) < 1e-4 * abs(c64)


@pytest.mark.parametrize("sgs", [0, 1])
@pytest.mark.parametrize("prec", [scl.DP, scl.SP])
def test_physical_tape_matches_transformed_state(prec, sgs, monkeypatch):
    """The backward with the physical-space tape of the recomputed forward
    (u, w [, S_ij] kept by the forward's x-pass) against the backward that
    transforms them from the stage states in every product."""
    n, nsteps, ckpt, cs2 = 32, 6, 4, 0.0225
    print(f"case adjoint physical tape prec={prec} sgs={sgs}", flush=True)
    k2 = _k2int(n)
    cd = scl._cdtype(prec)
    res = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("SCL_ADJ_PHYS_TAPE", mode)
        p, a, u0, tgt = _run_case(n, nsteps, ckpt, sgs, cs2, prec)
        _, gT = loss_grad(a.forward(u0, cs2), tgt, k2)
        res[mode] = a.backward([x.astype(cd) for x in gT])
        a.close()
        p.close()
    (g1, c1), (g0, c0) = res["1"], res["0"]
    err = max(np.abs(x - y).max() for x, y in zip(g1, g0)) / max(np.abs(y).max() for y in g0)
    tol = 1e-12 if prec == scl.DP else 2e-6
    print("tape vs no tape: grad rel err", err, "dcs2", c1, c0, flush=True)
    assert err < tol and abs(c1 - c0) <= tol * abs(c0)


@pytest.mark.parametrize("nsteps,ckpt", [(6, 6), (7, 3)])
def test_backward_after_vjp_and_repeated(nsteps, ckpt):
    """The forward leaves its last segment taped; a vjp call in between
    overwrites the tape and a second backward follows the first: every
    backward must give the gradient of a backward right after the forward."""
    n, cs2 = 32, 0.0225
    print(f"case adjoint backward order nsteps={nsteps} ckpt={ckpt}", flush=True)
    k2 = _k2int(n)
    p, a, u0, tgt = _run_case(n, nsteps, ckpt, 1, cs2, scl.DP)
    _, gT = loss_grad(a.forward(u0, cs2), tgt, k2)
    g_ref, c_ref = a.backward(gT)
    g_again, c_again = a.backward(gT)
    a.forward(u0, cs2)
    a.vjp(u0, _random_field(n, 7), cs2)
    g_vjp, c_vjp = a.backward(gT)
    # a vjp at another C_s^2 between the forward and the backward: the
    # backward still runs at the forward's
    a.forward(u0, cs2)
    a.vjp(u0, _random_field(n, 8), 4.0 * cs2)
    g_vjp2, c_vjp2 = a.backward(gT)
    for g, c in ((g_again, c_again), (g_vjp, c_vjp), (g_vjp2, c_vjp2)):
        err = max(np.abs(x - y).max() for x, y in zip(g, g_ref)) / max(np.abs(y).max() for y in g_ref)
        print("grad rel diff", err, "dcs2", c, c_ref, flush=True)
        assert err < 1e-13 and abs(c - c_ref) <= 1e-13 * abs(c_ref)
    a.close()
    p.close()


def test_torch_wrapper_matches_c_abi():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible to torch")
    from speccodelib_amd.torch_adjoint import rk4_map

    n, nsteps, ckpt, cs2 = 32, 6, 4, 0.0225
    print("case adjoint torch wrapper", flush=True)
    k2 = _k2int(n)
    p, a, u0, tgt = _run_case(n, nsteps, ckpt, 1, cs2, scl.DP)
    dev = torch.device("cuda", 0)
    U0 = torch.tensor(np.stack(u0), device=dev, requires_grad=True)
    Cs = torch.tensor(0.15, dtype=torch.float64, device=dev, requires_grad=True)
    T = torch.tensor(np.stack(tgt), device=dev)
    K2 = torch.tensor(k2, device=dev)
    H = torch.ones_like(K2)
    H[:, :, 0] = 0.5
    UT = rk4_map(U0, Cs * Cs, a)
    L = (H * ((UT - T).abs() ** 2 + K2 * UT.abs() ** 2)).sum()
    # a host forward of other data at another C_s^2 and a vjp in between:
    # the torch backward must still differentiate its own forward
    a.forward(tgt, 0.04)
    a.vjp(u0, _random_field(n, 9), 0.09)
    L.backward()
    uT = a.forward(u0, 0.15 * 0.15)
    L0, gT = loss_grad(uT, tgt, k2)
    g0, dcs2 = a.backward(gT)
    gt = U0.grad.cpu().numpy()
    err = max(np.abs(gt[q] - g0[q]).max() for q in range(3)) / max(np.abs(x).max() for x in g0)
    Lv = float(L.detach())
    print("torch loss", Lv, L0, "grad rel err", err, "dCs", float(Cs.grad), 2 * 0.15 * dcs2, flush=True)
    assert abs(Lv - L0) <= 1e-12 * L0
    assert err < 1e-12
    assert abs(float(Cs.grad) - 2 * 0.15 * dcs2) <= 1e-10 * abs(dcs2)
    a.close()
    p.close()


def test_adjoint_128_200_steps():
    """128^3 LES (C_s = 0.15), 200 steps of dt = 0.002, nu = 0.005, checkpoints every 20."""
    n, nsteps, ckpt, cs2 = 128, 200, 20, 0.15**2
    print("case adjoint 128^3 200 steps: fp64 fd + fp32 vs fp64", flush=True)
    p, a, u0, tgt, g64, c64 = _fd_check(n, nsteps, ckpt, 1, cs2, 8, 1e-4)
    k2 = _k2int(n)
    a.close()
    p.close()
    p, a, u0, tgt = _run_case(n, nsteps, ckpt, 1, cs2, scl.SP)
    uT = a.forward(u0, cs2)
    _, gT = loss_grad(uT, tgt, k2)
    g32, c32 = a.backward([x.astype(np.complex64) for x in gT])
    f, b = a.timing()
    err = max(np.abs(x - y).max() for x, y in zip(g32, g64)) / max(np.abs(y).max() for y in g64)
    print(f"fp32: forward {f:.1f} ms backward {b:.1f} ms ratio {b / f:.2f}; gradient vs fp64 rel err {err:.2e}; "
          f"dcs2 {c32:.6e} vs {c64:.6e}", flush=True)
    assert err < 1e-3 and abs(c32 - c64) < 1e-3 * abs(c64)
    a.close()
    p.close()
