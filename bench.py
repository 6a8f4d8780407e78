"""Throughput of the IF-RK4 forced-DNS step on one MI355X.

Workload: forced DNS on [0, 2pi)^3 at 512^3 in FP32 (dns-deform.f90 physics:
nu = 0.01, dt = 0.002, constant-power forcing on k^2 < 4 with eps = 0.1, 2/3
truncation), seeded k^4 exp(-2 k^2/16) initial field (iseed = 1, u0 = 1).
A step is one Lawson RK4 step: 4 RHS evaluations (5 FFT/pointwise passes
each) plus the stage updates, on fields resident in HBM.

Prints ONE JSON line.  `value` is steps/s of the timed region (K steps
bracketed by device synchronisation); per-kernel times come from HIP events
recorded on the solver's stream around every launch of that same region.
The roofline entry is for the kernel with the largest share of the step.
cpu_baseline times the oracle (the FP32 numpy/scipy restatement of the same
step) on the host's cores for one step at the same size.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK4 steps/s at 512³ FP32 (% HBM roofline); adjoint/forward time ratio"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def field_bytes(n):
    """F(N): one FP32 complex half-spectrum field, 8 (N/2+1) N^2 bytes."""
    return 8 * (n // 2 + 1) * n * n


# Algorithmic HBM bytes per launch, in units of F (one read or write of a
# field counts F).  The forced-mode coefficient kernel and the kx=0-plane
# Hermitian copy move O(N^2) bytes and count 0.
KIND_BYTES_F = {
    "force_coef": 0.0,
    "zinv_vort": 9.0,    # read u (3), write u and w (6)
    "y_inverse": 12.0,   # 6 fields read + written
    "x_pass": 9.0,       # read 6, write 3
    "y_forward": 6.0,    # 3 fields read + written
    "zfwd_rk": 13.5,     # stages 1..4: 12, 15, 15, 12 (W, v, du in; du, S or v out)
    "herm_fix": 0.0,
}
STEP_BYTES_F = 198.0  # 4 x (9 + 12 + 9 + 6) + (12 + 15 + 15 + 12)


def cpu_cores():
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
    return {
        "value": steps / dt,
        "unit": "steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{steps} IF-RK4 step(s) of the same {n}^3 FP32 forced DNS, oracle (numpy + scipy.fft "
        f"workers={cores}), {dt:.1f} s",
    }


def load_traffic(kernel, n):
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{n}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph-check", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the path's multi-GPU sharding "
              "is a design, not a build: see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import speccodelib_amd as scl

    n = a.n
    plan = scl.Plan3D(n, n, n, scl.SP)
    dns = scl.DNS(plan, nu=0.01, dt=0.002, forced=1, fkmax=2.0, eps=0.1, use_graph=True)
    dns.initialize(1, 4.0, 1.0)
    dns.step(a.warmup)
    dns.sync()
    # graph-replay timing (the production path)
    t0 = time.perf_counter()
    dns.step(a.steps)
    dns.sync()
    graph_ms = (time.perf_counter() - t0) * 1e3 / a.steps
    print(f"[bench] graph replay: {graph_ms:.3f} ms/step", file=sys.stderr, flush=True)

    # timed region: K steps with HIP events around every launch
    dns.sync()
    t0 = time.perf_counter()
    ev_total_ms, kinds = dns.step_timed(a.steps)
    dns.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3
    ms_per_step = wall_ms / a.steps
    print(f"[bench] timed region: {ms_per_step:.3f} ms/step (events {ev_total_ms / a.steps:.3f})", file=sys.stderr, flush=True)
    dg = dns.diagnostics()

    F = field_bytes(n)
    kern = {}
    for k, (ms, cnt) in kinds.items():
        if cnt == 0:
            continue
        avg_s = ms / cnt * 1e-3
        b = KIND_BYTES_F[k] * F
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
        "ns_per_point_step": ms_per_step * 1e6 / n**3,
        "graph_ms_per_step": graph_ms,
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
    if not a.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(n, a.cpu_steps, scl)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
