"""Benchmark: RK steps/s of the per-step renormalized DMFT update at L=512, FP64, one MI355X.

Workload: L=512 grid package (poly order 9, spline order 5, Tmax 1e5, alpha=delta=0),
p=3, q=4, lambda=0.5, Gamma=0.6, T0=inf (the production-size plain configuration).
The quench is first integrated from its initial condition for --prep steps (a real
history of T1 ~ prep rows, like the reference's [BENCH] harness runs on a live
state); then `steps` main-loop iterations are timed.  One step = one adaptive RK
step (RK54: 7 RHS evaluations with their interpolations, plus the dt rule)
through the C ABI (dmfe_loop_iteration); all inputs resident in HBM.

Prints one JSON line.  --gpus N with N > 1 is rejected: this build targets one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "RK steps/s at L=512 FP64; instance-steps/s batch-64 vs looped"
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes_rhs(N, m):
    """Bytes one rhs_rows_kernel launch must move (one RHS evaluation, one instance):
    24m N^2 stencil weights (A1, A2, B2) + 12 N^2 int32 stencil starts + 8 N^2 B2
    positions + 64 N^2 B1 history (4 row slices x (QK,QR)) + 64 N^2 B2 history
    (8 array-row scalars per target after window reuse) = (24m + 148) N^2."""
    return (24 * m + 148) * N * N


def make_grid(N=512):
    import dynamite_amd as dm
    t0 = time.time()
    g = dm.generate_grid(N, 1e5, 5, dm.INTERP_POLY, 9, 0)
    log(f"[bench] grid L={N} m={g.m} generated in {time.time() - t0:.1f}s")
    return g


PARAMS = dict(p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.6)


def gpu_run(g, prep, steps, warmup):
    import dynamite_amd as dm
    prm = dm.default_params(p=PARAMS["p"], p2=PARAMS["p2"], lambda_=PARAMS["lam"], T0=PARAMS["T0"],
                            Gamma=PARAMS["Gamma"])
    ctx = dm.Context(g, prm, rows_hint=prep + steps + warmup + 64)
    ctx.init_quench()
    t0 = time.perf_counter()
    n, _ = ctx.run(1e300, prep)
    ctx.synchronize()
    log(f"[bench] prepared a live state: {n} steps in {time.perf_counter() - t0:.2f}s, T1={ctx.T1}, "
        f"control={ctx.get_control()}")
    for _ in range(warmup):
        ctx.loop_iteration()
    ctx.synchronize()
    state, control = ctx.get_state(), ctx.get_control()
    dm.profile(ctx, True)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.loop_iteration()
    ctx.synchronize()
    el = time.perf_counter() - t0
    prof = {k: dm.profile_read(ctx, k) for k in (0, 1, 2)}
    dm.profile(ctx, False)
    T1 = ctx.T1
    ctx.close()
    return el, prof, state, control, T1


def cpu_baseline(g, state, control, budget_s=15.0):
    """The reference's own CPU step (built from its sources with its CPU flags) on a
    bounded sample of the same workload - the same live state, the same dt; falls back
    to the C oracle (one core) when the reference build is absent."""
    from oracle_lib import REF_FAST_PATH
    cores = min(16, os.cpu_count() or 1)
    kind = "reference"
    kw = dict(p=PARAMS["p"], p2=PARAMS["p2"], lam=PARAMS["lam"], T0=PARAMS["T0"], Gamma=PARAMS["Gamma"])
    if os.path.exists(REF_FAST_PATH):
        os.environ["OMP_NUM_THREADS"] = str(cores)
        os.environ.setdefault("OMP_PROC_BIND", "close")
        from oracle_lib import RefLib
        cpu = RefLib(g, path=REF_FAST_PATH, **kw)
    else:
        from oracle_lib import Oracle
        cpu = Oracle(g, **kw)
        kind, cores = "port", 1
    st = state
    cpu.set_history(st["t1grid"], st["dtratio"], st["rvec"], st["drvec"], st["QK"], st["QR"], st["dQK"], st["dQR"])
    cpu.interpolate(0)
    cpu.set_control(control["delta_t"], control["delta"], control["delta_old"], int(control["loop"]),
                    int(control["method"]))
    n = 0
    t0 = time.perf_counter()
    while True:
        cpu.loop_iteration()
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 500:
            break
    return dict(value=n / el, unit="RK steps/s", cores=cores, kind=kind,
                sample=f"{n} main-loop RK steps of the same L=512 live state (T1={len(st['t1grid'])}), "
                       f"{el:.1f}s wall")


def read_traffic():
    """HBM bytes per rhs_rows_kernel launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", "rhs_rows_L512_pmc.json")
    try:
        with open(path) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--prep", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a plain HBM copy of this many MiB (counter calibration) and exit")
    a = ap.parse_args()
    if a.copy_probe > 0:
        import dynamite_amd as dm
        nb = int(a.copy_probe * 2**20)
        ms = dm.copy_probe(nb, 10)
        log(f"[bench] copy probe: {nb} bytes read + {nb} written, {ms * 1e3:.1f} us/launch, "
            f"{2 * nb / ms / 1e6:.0f} GB/s")
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {a.gpus}); "
              "multi-GPU sharding is described in DESIGN.md and not run.", file=sys.stderr)
        sys.exit(2)
    N = 512
    g = make_grid(N)
    log(f"[bench] GPU: {a.prep} prep + {a.warmup} warmup + {a.steps} timed steps")
    el, prof, state, control, T1 = gpu_run(g, a.prep, a.steps, a.warmup)
    ms_step = 1e3 * el / a.steps
    rhs_ms, rhs_n = prof[1]
    avg_rhs_s = rhs_ms / 1e3 / max(rhs_n, 1)
    B = algorithmic_bytes_rhs(N, g.m)
    achieved = B / avg_rhs_s / 1e9
    log(f"[bench] {ms_step:.3f} ms/step; rhs_rows {rhs_n} launches avg {avg_rhs_s * 1e6:.1f} us "
        f"-> {achieved:.0f} GB/s algorithmic; stage (finish+search) avg "
        f"{prof[2][0] / max(prof[2][1], 1) * 1e3:.1f} us x {prof[2][1]}; search-only avg "
        f"{prof[0][0] / max(prof[0][1], 1) * 1e3:.1f} us x {prof[0][1]}")
    traffic = read_traffic()
    out = {
        "metric": METRIC,
        "value": a.steps / el,
        "unit": "RK steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain quench, L=512 poly-9 grid (spline 5, Tmax 1e5), p=3 q=4 lambda=0.5 "
                               "Gamma=0.6 T0=inf; adaptive RK steps on the live state after the prep steps",
                   "L": N, "stencil_width": g.m, "prep_steps": a.prep, "T1_end": T1,
                   "method": "RK54" if control["method"] == 1 else "SSPRK104", "delta_t": control["delta_t"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "rhs_rows_kernel", "bytes_per_launch": B,
                     "avg_launch_us": avg_rhs_s * 1e6, "launches": rhs_n},
    }
    if not a.no_cpu:
        log(f"[bench] CPU baseline (budget {a.cpu_budget:.0f}s)")
        out["cpu_baseline"] = cpu_baseline(g, state, control, a.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
