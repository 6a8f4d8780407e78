"""Throughput of the IF-RK4 forced-DNS step on one MI355X.

Workload: forced DNS on [0, 2pi)^3 at 512^3 in FP32 (dns-deform.f90 physics:
nu = 0.01, dt = 0.002, constant-power forcing on k^2 < 4 with eps = 0.1, 2/3
truncation), seeded k^4 exp(-2 k^2/16) initial field (iseed = 1, u0 = 1).
A step is one Lawson RK4 step: 4 RHS evaluations (5 FFT/pointwise passes
each) plus the stage updates, on fields resident in HBM.

Prints ONE JSON line.  `value` is steps/s of the timed region: K steps of
graph replay between device synchronisations.  The per-kernel times come
from HIP events recorded on the solver's stream around every launch of the
same K steps run again with direct launches.
The roofline entry is for the kernel with the largest share of the step.
cpu_baseline times the oracle (the FP32 numpy/scipy restatement of the same
step) on the host's cores for one step at the same size.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK4 steps/s at 512³ FP32 (% HBM roofline); adjoint/forward time ratio"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def field_bytes(n):
    """F(N): one FP32 complex half-spectrum field, 8 (N/2+1) N^2 bytes."""
    return 8 * (n // 2 + 1) * n * n


# Algorithmic HBM bytes per launch of these kernels, in units of F (one read
# or write of a field counts F).  The forced-mode coefficient kernel and the
# kx=0-plane Hermitian copy move O(N^2) bytes and count 0.
KIND_BYTES_F = {
    "force_coef": 0.0,
    "zinv_vort": 8.0,    # read u (3), write u, w_x, w_y (5)
    "y_inverse": 11.0,   # read 5, write 6 (w_z formed in the pass)
    "x_pass": 9.0,       # read 6, write 3
    "y_forward": 6.0,    # 3 fields read + written
    "zfwd_rk": 12.75,    # stages 1..4: 12, 15, 15, 9 (W, v, du in; du, S or v out)
    "herm_fix": 0.0,
    # z-forward + stage + z-inverse of the next stage state (stage chain):
    # stages 1..3: 14, 17, 17 (W, v, du in; du and W[0..4] out)
    "zfwd_zinv": 16.0,
    # stage 4 + z-inverse of the new v for the next step: W and du in (3 + 3),
    # v and W[0..4] out (3 + 5)
    "zfwd4_zinv": 14.0,
}
# The step model that the roofline fraction is priced against: one read +
# write per axis pass (z-inverse 9, y-inverse 12, x-pass 9, y-forward 6;
# z-forward + stage 12, 15, 15, 12).
STEP_BYTES_F = 198.0
# What these kernels move per step: 4 x (11 + 9 + 6) for the y-inverse,
# x-pass and y-forward, and the four z-forwards, each fused with the next
# z-inverse (14 + 17 + 17, and 14 for stage 4 with the next step's): 166 F.
IMPL_STEP_BYTES_F = 166.0


def kind_bytes_f(kinds):
    """Per-launch bytes of each kind in one step_timed run: with stage
    chaining the unfused z-forward is stage 4 only."""
    out = dict(KIND_BYTES_F)
    if kinds.get("zfwd_zinv", (0.0, 0))[1]:
        out["zfwd_zinv"] = 16.0  # (14 + 17 + 17) / 3
        out["zfwd_rk"] = 9.0
    return out


def cpu_cores():
    """The threads the baseline may use: the host's cores available to this
    process, at most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(n, steps, scl):
    from oracle import spectral3d as orc

    cores = cpu_cores()
    orc.set_fft_backend("scipy", workers=cores)
    v = scl.initialize(n, n, n, 1, 4.0, 1.0, scl.SP)  # input data only
    d = orc.DNS(orc.DNSParams(nx=n, nu=0.01, dt=0.002, forced=True, fkmax=2.0, eps=0.1, prec="sp"))
    t0 = time.perf_counter()
    for _ in range(steps):
        v = d.rk4_step(*v)
    dt = time.perf_counter() - t0
    # one thread, on the same configuration at half the size (a bounded sample)
    n1 = max(16, n // 2)
    orc.set_fft_backend("scipy", workers=1)
    v1 = scl.initialize(n1, n1, n1, 1, 4.0, 1.0, scl.SP)
    d1 = orc.DNS(orc.DNSParams(nx=n1, nu=0.01, dt=0.002, forced=True, fkmax=2.0, eps=0.1, prec="sp"))
    t0 = time.perf_counter()
    d1.rk4_step(*v1)
    dt1 = time.perf_counter() - t0
    return {
        "value": steps / dt,
        "unit": "steps/s",
        "cores": cores,
        "host_cores": os.cpu_count(),
        "kind": "port",
        "sample": f"{steps} IF-RK4 step(s) of the same {n}^3 FP32 forced DNS, oracle (numpy + scipy.fft "
        f"workers={cores}), {dt:.1f} s",
        "ns_per_point_step": dt * 1e9 / (steps * n**3),
        "one_thread": {
            "value": 1.0 / dt1,
            "unit": "steps/s",
            "ns_per_point_step": dt1 * 1e9 / n1**3,
            "sample": f"1 IF-RK4 step of the same forced DNS at {n1}^3, oracle with scipy.fft workers=1, "
            f"{dt1:.1f} s",
        },
    }


# Uneven 2D vorticity ensembles: 192 x 256^2, 48 x 512^2, 16 x 1024^2, FP32,
# nu = 1e-5 k^8 hyperviscosity, linear damping 0.1, forcing on 4 <= |k| <= 5,
# dt = 5e-4, 1000 IF-RK4 steps.  Per RHS 15 F2 (y-inverse 1 in + 3 out,
# x-pass 3 + 2, y-forward 2 + state and stage I/O), 58 F2 per step with
# F2 = 8 (n/2 + 1) n bytes per member (the model the roofline fraction is
# priced against; the stage chain moves 51 F2 per step).
ENS2D_GROUPS = [(256, 192), (512, 48), (1024, 16)]
ENS2D_KW = dict(dt=5e-4, c_hyper=1e-5, n_hyper=8, damp=0.1, fkmin=4.0, fkmax=5.0, eps=0.01)
ENS2D_STEP_F2 = 58.0


def bench_ens2d(scl, steps, member_steps):
    f2 = sum(m * 8 * (n // 2 + 1) * n for n, m in ENS2D_GROUPS)
    eb = scl.Ensemble2D(ENS2D_GROUPS, prec=scl.SP, **ENS2D_KW)
    eb.initialize(seed0=-1, k0=8.0, energy=0.5)
    best = eb.tune(reps=3)
    sched = [eb.get_schedule(g) for g in range(len(ENS2D_GROUPS))]
    eb.step(5)
    ms = eb.step_timed(steps)
    d = eb.diagnostics()
    eb.close()
    ep = scl.Ensemble2D(ENS2D_GROUPS, prec=scl.SP, per_member=True, **ENS2D_KW)
    ep.initialize(seed0=-1, k0=8.0, energy=0.5)
    for g, sc in enumerate(sched):
        ep.set_schedule(g, *sc)
    ep.step(2)
    msp = ep.step_timed(member_steps)
    ep.close()
    bms, pms = ms / steps, msp / member_steps
    step_bytes = ENS2D_STEP_F2 * f2
    finite = bool(np.all(np.isfinite(d)))
    print(f"[bench] ens2d: batched {bms:.3f} ms/step, per-member {pms:.3f} ms/step", file=sys.stderr, flush=True)
    return {
        "workload": "2D vorticity ensembles 192x256^2 + 48x512^2 + 16x1024^2 FP32, IF-RK4 dt=5e-4, "
                    "1e-5 k^8 hyperviscosity, damping 0.1, forcing 4<=|k|<=5 eps=0.01, IC k0=8 energy 0.5",
        "batched_steps": steps,
        "batched_ms_per_step": bms,
        "batched_steps_per_s": 1e3 / bms,
        "per_member_steps": member_steps,
        "per_member_ms_per_step": pms,
        "speedup_vs_per_member": pms / bms,
        "step_alg_bytes": step_bytes,
        "hbm_roofline_frac_step": step_bytes / (bms * 1e-3) / HBM_PEAK,
        "ns_per_point_step": bms * 1e6 / sum(m * n * n for n, m in ENS2D_GROUPS),
        "schedules_e_cw_rowgrid": sched,
        "tuned_ms_per_step": [float(x) for x in best],
        "finite": finite,
        "energy_mean": float(d[:, 0].mean()),
    }


# Discrete adjoint: 128^3 static Smagorinsky LES (C_s = 0.15, nu = 0.005,
# dt = 0.002), 200 IF-RK4 steps with checkpoints every 20, loss
# sum' h (|u(T) - u_tgt|^2 + k^2 |u(T)|^2) (h = 1/2 on kx = 0), target from
# seed -2 evolved the same way.  ratio = backward time (recomputation
# included) / forward time.
def bench_adjoint(scl, n=128, nsteps=200, ckpt=20, cs=0.15):
    p = scl.Plan3D(n, n, n, scl.SP)
    a = scl.Adjoint(p, nsteps, ckpt, nu=0.005, dt=0.002, sgs=1)
    u0 = scl.initialize(n, n, n, -1, 3.0, 1.0, scl.SP)
    tgt = a.forward(scl.initialize(n, n, n, -2, 3.0, 1.0, scl.SP), cs * cs)
    uT = a.forward(u0, cs * cs)
    kx, ky, kz = (np.arange(n // 2 + 1), np.fft.fftfreq(n, 1.0 / n), np.fft.fftfreq(n, 1.0 / n))
    k2 = kx[None, None, :] ** 2 + ky[None, :, None] ** 2 + kz[:, None, None] ** 2
    h = np.ones(k2.shape)
    h[:, :, 0] = 0.5
    gT = [(2 * h * (u - t) + 2 * h * k2 * u).astype(np.complex64) for u, t in zip(uT, tgt)]
    loss = float(sum((h * (np.abs(u - t) ** 2 + k2 * np.abs(u) ** 2)).sum() for u, t in zip(uT, tgt)))
    g0, dcs2 = a.backward(gT)
    f, b = a.timing()
    a.close()
    p.close()
    print(f"[bench] adjoint {n}^3: forward {f:.1f} ms, backward {b:.1f} ms, ratio {b / f:.2f}", file=sys.stderr,
          flush=True)
    return {
        "workload": f"adjoint of {nsteps} IF-RK4 steps, {n}^3 FP32 static Smagorinsky LES C_s={cs}, nu=0.005, "
                    f"dt=0.002, checkpoints every {ckpt}",
        "forward_ms": f,
        "backward_ms": b,
        "ratio": b / f,
        "loss": loss,
        "dL_dCs": 2 * cs * dcs2,
        "grad_norm": float(np.sqrt(sum((np.abs(x) ** 2).sum() for x in g0))),
    }


def copy_calibration(nbytes):
    """One fill and one device-to-device copy of nbytes through the HIP runtime
    that libscl.so loaded, so that a counter profile of this process holds two
    kernels of known traffic (fill: nbytes written; copy: nbytes read and
    written)."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so.7")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(nbytes)
    if hip.hipMalloc(ctypes.byref(src), n) or hip.hipMalloc(ctypes.byref(dst), n):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, n)
    hip.hipMemcpy(dst, src, n, 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"[bench] copy calibration: fill + copy of {nbytes} bytes", file=sys.stderr, flush=True)


# 1024^3 FP32 DNS with helical forcing (force_heli.f90): nu = 2.5e-4,
# dt = 2.5e-4, forcing on k^2 < 4 with eps = 0.1 and eta = 2 k_f eps
# (relative helicity 1 at k_f = 2), seeded k^4 exp(-2 (k/4)^2) field.
def bench_heli1024(scl, steps, n=1024):
    plan = scl.Plan3D(n, n, n, scl.SP)
    d = scl.DNS(plan, nu=2.5e-4, dt=2.5e-4, forced=2, fkmax=2.0, eps=0.1, eta=0.4, use_graph=True)
    t0 = time.perf_counter()
    d.initialize(1, 4.0, 1.0)
    d.sync()
    ic_s = time.perf_counter() - t0
    d.step(1)
    d.sync()
    t0 = time.perf_counter()
    d.step(steps)
    d.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    _, kinds = d.step_timed(2)
    dg = d.diagnostics()
    ek, hk = d.spectrum(), d.helicity_spectrum()
    d.close()
    plan.close()
    F = field_bytes(n)
    print(f"[bench] helical {n}^3: {ms:.1f} ms/step", file=sys.stderr, flush=True)
    return {
        "workload": f"helical forced DNS {n}^3 FP32, nu=2.5e-4, dt=2.5e-4, force_heli k^2<4 eps=0.1 eta=0.4",
        "steps": steps,
        "ms_per_step": ms,
        "steps_per_s": 1e3 / ms,
        "hbm_roofline_frac_step": STEP_BYTES_F * F / (ms * 1e-3) / HBM_PEAK,
        "kernel_avg_us": {k: (v[0] / v[1] * 1e3 if v[1] else None) for k, v in kinds.items()},
        "initial_field_s": ic_s,
        "energy": dg["energy"],
        "helicity": dg["helicity"],
        "Ek_1_8": [float(x) for x in ek[:8]],
        "Hk_1_8": [float(x) for x in hk[:8]],
    }


# 256^3 FP32 static-Smagorinsky LES with a passive scalar (spectral_les.f90,
# spectral_dns_scalar.f90): C_s = 0.17, Delta = 2 pi/256, nu = 1e-4,
# dt = 0.001, forcing on |k| <= 2.5, scalar with unit mean gradient and
# Sc = 0.7 started from zero; seeded k^4 exp(-2 (k/4)^2) field.
def bench_les256(scl, steps, n=256):
    plan = scl.Plan3D(n, n, n, scl.SP)
    d = scl.DNS(plan, nu=1e-4, dt=0.001, forced=1, fkmax=2.5, eps=0.1, sgs=1, cs2=0.17**2, delta=2 * np.pi / n,
                scalar=True, grad=1.0, schmidt=0.7, use_graph=True)
    d.initialize(1, 4.0, 1.0)
    d.set_scalar(np.zeros((n, n, n // 2 + 1), np.complex64))
    d.step(2)
    d.sync()
    t0 = time.perf_counter()
    d.step(steps)
    d.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    dg = d.diagnostics()
    d.close()
    plan.close()
    print(f"[bench] LES + scalar {n}^3: {ms:.2f} ms/step", file=sys.stderr, flush=True)
    return {
        "workload": f"static Smagorinsky LES + passive scalar {n}^3 FP32, C_s=0.17, nu=1e-4, dt=0.001, Sc=0.7, "
                    "forcing |k|<=2.5",
        "steps": steps,
        "ms_per_step": ms,
        "steps_per_s": 1e3 / ms,
        "ns_per_point_step": ms * 1e6 / n**3,
        "energy": dg["energy"],
        "scalar_variance": dg["scalar_variance"],
        "nonfinite": dg["nonfinite"],
    }


def bench_ab2(scl, steps, n=512):
    """The production drivers' IF-AB2 stepper (spectral_dns_dealiased.f90: CFL
    dt, umax refreshed every 10 steps) on the 512^3 FP32 forced DNS: one RHS
    per step on the same passes, the steady step as a graph."""
    plan = scl.Plan3D(n, n, n, scl.SP)
    d = scl.DNS(plan, nu=0.01, forced=1, fkmax=2.0, eps=0.1, stepper=1, beta=0.5, dtw=0.1, ieout=10,
                umax=3.0 * np.sqrt(3.0), kcut=n // 2 - 1.0, use_graph=True)
    d.initialize(1, 4.0, 1.0)
    d.step(4)
    d.sync()
    t0 = time.perf_counter()
    d.step(steps)
    d.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    t, dt_last, k = d.step_info()
    dg = d.diagnostics()
    d.close()
    plan.close()
    print(f"[bench] IF-AB2 {n}^3: {ms:.2f} ms/step", file=sys.stderr, flush=True)
    return {
        "workload": f"IF-AB2 (CFL dt, beta 0.5, umax every 10 steps) forced DNS {n}^3 FP32, nu=0.01",
        "steps": steps,
        "ms_per_step": ms,
        "steps_per_s": 1e3 / ms,
        "time": t,
        "dt_last": dt_last,
        "energy": dg["energy"],
    }


# turb_2d.f90's scheme on the configs[2] groups: per step an unshifted and a
# shifted half RHS, 28 F2 (y-inverse 1 + 3 twice, x-pass 3 + 2 twice,
# y-forward 2 + 1 then 2 + 1 + state and owz in and out).
TURB2D_STEP_F2 = 28


def bench_turb2d(scl, steps):
    f2 = sum(m * 8 * (n // 2 + 1) * n for n, m in ENS2D_GROUPS)
    e = scl.Ensemble2D(ENS2D_GROUPS, prec=scl.SP, dt=1e-4, scheme=1, sigma=1.0 / 400.0, nu=0.5, eta=5.0, kf=4.0)
    e.initialize(seed0=-1, k0=8.0, energy=0.5)
    e.tune(reps=3)
    e.step(4)
    ms = e.step_timed(steps) / steps
    d = e.diagnostics()
    e.close()
    print(f"[bench] turb_2d scheme: {ms:.3f} ms/step", file=sys.stderr, flush=True)
    return {
        "workload": "turb_2d.f90 step (phase-shift dealiasing, (k sigma)^16 + (k/nu)^-4, band forcing, IF-AB2) on "
                    "192x256^2 + 48x512^2 + 16x1024^2 FP32, dt=1e-4, sigma=1/400, nu=0.5, eta=5, kf=4",
        "steps": steps,
        "ms_per_step": ms,
        "steps_per_s": 1e3 / ms,
        "step_alg_bytes": TURB2D_STEP_F2 * f2,
        "hbm_roofline_frac_step": TURB2D_STEP_F2 * f2 / (ms * 1e-3) / HBM_PEAK,
        "finite": bool(np.all(np.isfinite(d))),
    }


def bench_padded(scl, reps, n=256):
    """The 3/2-rule padded Lamb vector (scl_convec_padded) at 256^3 FP32 on
    host arrays: six fields in, three out across PCIe per call, the products
    on a 512^3 plan."""
    plan = scl.Plan3D(n, n, n, scl.SP)
    rng = np.random.default_rng(0)
    f = [(rng.standard_normal(plan.shape) + 1j * rng.standard_normal(plan.shape)).astype(np.complex64)
         for _ in range(6)]
    plan.convec_padded(*[a.copy() for a in f])  # creates the padded plan
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.convec_padded(*f)
    ms = (time.perf_counter() - t0) * 1e3 / reps
    plan.close()
    print(f"[bench] padded convec {n}^3: {ms:.1f} ms/call", file=sys.stderr, flush=True)
    return {"workload": f"scl_convec_padded {n}^3 FP32 (2n padded grid), host arrays", "calls": reps,
            "ms_per_call": ms}


def bench_stream(scl, nbytes=2 << 30, reps=10):
    """The HBM rate plain streaming kernels reach (the practical ceiling of
    the passes): read, write and copy of 2 GiB buffers."""
    out = {"bytes": nbytes, "reps": reps}
    for kind, name in ((0, "read"), (1, "write"), (2, "copy")):
        ms, gbps = scl.stream_probe(kind, nbytes, reps)
        out[name + "_GBps"] = gbps
        out[name + "_ms"] = ms
    print(f"[bench] stream probe: read {out['read_GBps']:.0f} GB/s, write {out['write_GBps']:.0f} GB/s, "
          f"copy {out['copy_GBps']:.0f} GB/s", file=sys.stderr, flush=True)
    return out


def load_traffic(kernel, n):
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{n}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def bench_main(scl, a):
    """The 512^3 forced DNS: the headline line (everything but the extra sections)."""
    n = a.n
    plan = scl.Plan3D(n, n, n, scl.SP)
    dns = scl.DNS(plan, nu=0.01, dt=0.002, forced=1, fkmax=2.0, eps=0.1, use_graph=not a.eager)
    dns.initialize(1, 4.0, 1.0)
    dns.step(a.warmup)
    dns.sync()
    # timed region: K steps of graph replay (the production path) between
    # device synchronisations
    t0 = time.perf_counter()
    dns.step(a.steps)
    dns.sync()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / a.steps
    print(f"[bench] timed region (graph replay): {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)

    # the same K steps again with a HIP event around every launch on the
    # solver's stream (direct launches: events inside a graph cannot be
    # timed here), for the per-kernel times
    dns.sync()
    ev_total_ms, kinds = dns.step_timed(a.steps)
    dns.sync()
    inst_ms = ev_total_ms / a.steps
    print(f"[bench] instrumented steps: {inst_ms:.3f} ms/step", file=sys.stderr, flush=True)
    dg = dns.diagnostics()
    if a.copy_calib:
        copy_calibration(a.copy_calib)

    F = field_bytes(n)
    kb = kind_bytes_f(kinds)
    kern = {}
    for k, (ms, cnt) in kinds.items():
        if cnt == 0:
            continue
        avg_s = ms / cnt * 1e-3
        b = kb[k] * F
        kern[k] = {"ms_per_step": ms / a.steps, "launches": cnt, "avg_us": avg_s * 1e6,
                   "alg_bytes_per_launch": b, "achieved_GBps": (b / avg_s / 1e9) if b else None}
    dom = max(kern, key=lambda k: kern[k]["ms_per_step"])
    dk = kern[dom]
    achieved = dk["alg_bytes_per_launch"] / (dk["avg_us"] * 1e-6)
    step_bytes = STEP_BYTES_F * F
    out = {
        "metric": METRIC,
        "value": 1e3 / ms_per_step,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"forced DNS {n}^3 FP32, Lawson IF-RK4 (dns-deform.f90), nu=0.01, dt=0.002, "
                   "force k^2<4 eps=0.1, 2/3 cut, IC k0=4 u0=1 iseed=1", "n": n, "parallelism": "single GPU"},
        "hbm_roofline_frac_step": step_bytes / (ms_per_step * 1e-3) / HBM_PEAK,
        "step_alg_bytes": step_bytes,
        "impl_step_bytes": IMPL_STEP_BYTES_F * F,
        "impl_bandwidth_frac_step": IMPL_STEP_BYTES_F * F / (ms_per_step * 1e-3) / HBM_PEAK,
        "ns_per_point_step": ms_per_step * 1e6 / n**3,
        "instrumented_ms_per_step": inst_ms,
        "adjoint_forward_ratio": None,
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": load_traffic(dom, n)},
        "kernels": kern,
        "check": {"energy": dg["energy"], "divmax": dg["divmax"], "nonfinite": dg["nonfinite"]},
    }
    if dg["nonfinite"]:
        print("[bench] non-finite values in the solution", file=sys.stderr)
        sys.exit(3)
    del dns
    plan.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph-check", action="store_true")
    ap.add_argument("--ens2d-steps", type=int, default=1000)
    ap.add_argument("--ens2d-member-steps", type=int, default=20)
    ap.add_argument("--no-ens2d", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--heli1024", type=int, default=5,
                    help="also time this many steps of the 1024^3 helical DNS (0: skip)")
    ap.add_argument("--les256", type=int, default=50,
                    help="also time this many steps of the 256^3 LES + scalar (0: skip)")
    ap.add_argument("--ab2", type=int, default=20, help="also time this many IF-AB2 steps at 512^3 (0: skip)")
    ap.add_argument("--turb2d", type=int, default=200,
                    help="also time this many steps of the turb_2d scheme on the 2D groups (0: skip)")
    ap.add_argument("--padded", type=int, default=3, help="also time this many padded convec calls (0: skip)")
    ap.add_argument("--no-main", action="store_true", help="skip the 512^3 DNS (profiling runs of the others)")
    ap.add_argument("--no-stream", action="store_true", help="skip the streaming HBM probe")
    ap.add_argument("--eager", action="store_true", help="launch the kernels directly instead of as graphs")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="bytes of a device-to-device copy and a fill run after the timed region (counter "
                         "calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU sharding of this "
              "path is a design, not a build", file=sys.stderr)
        sys.exit(2)

    import speccodelib_amd as scl

    if a.no_main:
        out = {"metric": METRIC, "value": None, "unit": "steps/s", "note": "512^3 DNS skipped (--no-main)"}
    else:
        out = bench_main(scl, a)
    out["hbm_stream"] = None if a.no_stream else bench_stream(scl)
    if out["hbm_stream"] and out.get("roofline"):
        # the dominant kernel's rate against what a plain copy reaches here
        out["roofline"]["frac_of_copy"] = out["roofline"]["achieved"] / out["hbm_stream"]["copy_GBps"]
    out["ens2d"] = None if a.no_ens2d else bench_ens2d(scl, a.ens2d_steps, a.ens2d_member_steps)
    if not a.no_adjoint:
        out["adjoint"] = bench_adjoint(scl)
        out["adjoint_forward_ratio"] = out["adjoint"]["ratio"]
    if a.heli1024:
        out["heli1024"] = bench_heli1024(scl, a.heli1024)
    if a.les256:
        out["les256"] = bench_les256(scl, a.les256)
    if a.ab2:
        out["ab2_512"] = bench_ab2(scl, a.ab2)
    if a.turb2d:
        out["turb2d"] = bench_turb2d(scl, a.turb2d)
    if a.padded:
        out["padded256"] = bench_padded(scl, a.padded)
    if not a.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(a.n, a.cpu_steps, scl)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
