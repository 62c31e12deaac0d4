"""Benchmark: RK steps/s of the per-step renormalized DMFT update at L=512, FP64, one MI355X.

Workload: L=512 grid package (poly order 9, spline order 5, Tmax 1e5, alpha=delta=0),
p=3, q=4, lambda=0.5, Gamma=0.6, T0=inf (the production-size plain configuration).
The quench is first integrated from its initial condition for --prep steps (a real
history of T1 ~ prep rows, like the reference's [BENCH] harness runs on a live
state); then `steps` main-loop iterations are timed.  One step = one adaptive RK
step (RK54: 7 RHS evaluations with their interpolations, plus the dt rule)
through the C ABI (dmfe_run, the main loop with its step control on the device; the
host-driven dmfe_loop_iteration is reported beside it); all inputs resident in HBM.

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


_GRIDS = {}


def make_grid(N=512):
    import dynamite_amd as dm
    if N in _GRIDS:
        return _GRIDS[N]
    t0 = time.time()
    g = dm.generate_grid(N, 1e5, 5, dm.INTERP_POLY, 9, 0)
    log(f"[bench] grid L={N} m={g.m} generated in {time.time() - t0:.1f}s")
    _GRIDS[N] = g
    return g


PARAMS = dict(p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.6)


USE_GRAPHS = True  # --no-graphs: the same launches issued directly (for tracers that cannot follow graphs)


def new_context(g, prm, rows_hint):
    import dynamite_amd as dm
    ctx = dm.Context(g, prm, rows_hint=rows_hint)
    if not USE_GRAPHS:
        dm.set_loop_graphs(ctx, False)
    return ctx


def gpu_run(g, prep, steps, warmup):
    """The plain quench at L=512: `prep` steps from the quench, then `warmup` + `steps`
    timed steps of dmfe_run (the main loop with the step control on the device, replayed
    hipGraphs); then 50 steps of the host-driven loop (dmfe_loop_iteration) with per-launch
    events for the in-step kernel averages, and back-to-back RHS launches."""
    import dynamite_amd as dm
    prm = dm.default_params(p=PARAMS["p"], p2=PARAMS["p2"], lambda_=PARAMS["lam"], T0=PARAMS["T0"],
                            Gamma=PARAMS["Gamma"])
    ctx = new_context(g, prm, rows_hint=prep + steps + warmup + 128)
    ctx.init_quench()
    t0 = time.perf_counter()
    n, _ = ctx.run(1e300, prep)
    ctx.synchronize()
    log(f"[bench] prepared a live state: {n} steps in {time.perf_counter() - t0:.2f}s, T1={ctx.T1}, "
        f"control={ctx.get_control()}")
    # the autotuner picks the RHS kernel shape (lanes per B2 entry, column chunks) on
    # this live history and installs it for the L=512 class
    best, table = dm.autotune(ctx, reps=20)
    tune = dict(shape={"G": best[0], "split": best[1]},
                candidates_us={f"G{G}_split{sp}": round(1e3 * ms, 2) for (G, sp), ms in table.items()})
    log(f"[bench] autotune: best G={best[0]} split={best[1]}; {tune['candidates_us']}")
    loop = ctx.get_control()["loop"]
    ctx.run(1e300, loop + warmup)
    ctx.synchronize()
    state, control = ctx.get_state(), ctx.get_control()
    loop = control["loop"]
    t0 = time.perf_counter()
    n, _ = ctx.run(1e300, loop + steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    if n != steps:
        raise RuntimeError(f"dmfe_run stopped after {n} of {steps} steps")
    # host-driven loop on the continuing state: per-launch events inside the steps
    dm.profile(ctx, True)
    host_steps = 50
    t0 = time.perf_counter()
    for _ in range(host_steps):
        ctx.loop_iteration()
    ctx.synchronize()
    el_host = time.perf_counter() - t0
    prof = {k: dm.profile_read(ctx, k) for k in (0, 1, 2)}
    dm.profile(ctx, False)
    # the RHS kernel alone, back to back on the final state (per-launch events inside
    # the steps also count the gap to the preceding stage launch)
    rhs_b2b_ms = dm.time_rhs(ctx, 200)
    T1 = ctx.T1
    ctx.close()
    host = dict(steps=host_steps, value=host_steps / el_host, ms_per_step=1e3 * el_host / host_steps)
    return el, prof, state, control, T1, rhs_b2b_ms, tune, host


def rhs_only_run(g, prep, reps):
    """The live L=512 state of the plain workload (`prep` dmfe_run steps from the quench),
    then `reps` back-to-back RHS launches of the installed shape on it: the dispatches
    that counter runs select by kernel name."""
    import dynamite_amd as dm
    prm = dm.default_params(p=PARAMS["p"], p2=PARAMS["p2"], lambda_=PARAMS["lam"], T0=PARAMS["T0"],
                            Gamma=PARAMS["Gamma"])
    ctx = new_context(g, prm, rows_hint=prep + 128)
    ctx.init_quench()
    n, _ = ctx.run(1e300, prep)
    ctx.synchronize()
    ms = dm.time_rhs(ctx, reps)
    out = dict(prep_steps=n, T1=ctx.T1, reps=reps, rhs_us=1e3 * ms,
               achieved_GBs=algorithmic_bytes_rhs(g.N, g.m) / (ms * 1e-3) / 1e9)
    ctx.close()
    log(f"[bench] rhs only: {out}")
    return out


def full_run(g, tmax, segment=10000, wall_limit=900.0):
    """The plain L=512 configuration integrated from the quench towards t = tmax with
    dmfe_run (device-side step control, the every-1e5-loops maintenance included), in
    segments of `segment` loops: per segment the loop count, t, dt, integrator, T1, E(t)
    and QK(t,0); T1 before and after each maintenance.  Stops at tmax, when the main loop
    stops by itself (delta_t < delta_t_min) or after wall_limit seconds."""
    import dynamite_amd as dm
    prm = dm.default_params(p=PARAMS["p"], p2=PARAMS["p2"], lambda_=PARAMS["lam"], T0=PARAMS["T0"],
                            Gamma=PARAMS["Gamma"])
    ctx = new_context(g, prm, rows_hint=segment + 256)
    ctx.init_quench()
    recs, maint = [], []
    t0 = time.perf_counter()
    loop, t, stop = 0, 0.0, "tmax"
    while True:
        T1_before = ctx.T1
        # the correlation records of the segment (t, E(t), QK(t,0) per loop; the last one
        # is the segment's end)
        n, corr = ctx.run(tmax, loop + segment, corr_cap=segment)
        ctx.synchronize()
        c = ctx.get_control()
        loop = int(c["loop"])
        if n == 0:
            stop = "delta_t < delta_t_min" if c["delta_t"] < 1e-5 else "loop stopped"
            break
        t, E, qk0 = (float(v) for v in corr[-1])
        el = time.perf_counter() - t0
        rec = dict(loop=loop, t=t, dt=c["delta_t"], method="RK54" if c["method"] == 1 else "SSPRK104",
                   T1=ctx.T1, E=E, QK_t0=qk0, wall_s=el)
        recs.append(rec)
        # the maintenance ran if a multiple of 1e5 loops was crossed in this segment
        if loop // 100000 > (loop - n) // 100000:
            maint.append(dict(loop=(loop // 100000) * 100000, T1_segment_start=T1_before, T1_segment_end=ctx.T1))
        log(f"[bench] full run: loop {loop} t {t:.6g} dt {c['delta_t']:.4g} {rec['method']} T1 {ctx.T1} "
            f"E {E:.10g} QK(t,0) {qk0:.10g} ({el:.0f} s)")
        if t >= tmax:
            break
        if n < segment:
            stop = "delta_t < delta_t_min" if c["delta_t"] < 1e-5 else "loop stopped"
            break
        if el > wall_limit:
            stop = "wall limit"
            break
    ctx.close()
    return dict(workload="plain quench, L=512 poly-9 grid (spline 5, Tmax 1e5), p=3 q=4 lambda=0.5 Gamma=0.6 "
                         "T0=inf, dmfe_run from the quench", tmax=tmax, stop=stop, loops=loop, t_final=t,
                wall_s=time.perf_counter() - t0, E_final=recs[-1]["E"], QK_t0_final=recs[-1]["QK_t0"],
                maintenance=maint, segments=recs)


def host_cpus():
    """CPUs this process may use, the machine's logical CPUs and its physical cores
    (distinct (package, core) pairs of /proc/cpuinfo among the allowed CPUs)."""
    allowed = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    phys, cpu, pkg = set(), None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "processor":
                    cpu, pkg = int(v), None
                elif k == "physical id":
                    pkg = v.strip()
                elif k == "core id" and cpu in allowed:
                    phys.add((pkg, v.strip()))
    except OSError:
        pass
    return dict(nproc=os.cpu_count(), allowed=len(allowed), physical_cores=len(phys) or None)


def _omp_threads(n):
    """Sets the OpenMP thread count of the calling thread (libgomp, which the reference
    build links) for the parallel regions it starts from now on."""
    import ctypes
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def cpu_baseline(g, state, control, budget_s=15.0):
    """The reference's own CPU step (built from its sources with its CPU flags) on a
    bounded sample of the same workload - the same live state, the same dt - with one
    OpenMP thread and with all physical cores (at most 16, the host share of one GPU),
    plus the wall time of the CPU-sized configuration run to its end (L=128 quench,
    Gamma=0.8, tmax 1e3) on both threads counts and on the GPU (dmfe_run).  Falls back to
    the C oracle (one core) when the reference build is absent."""
    from oracle_lib import REF_FAST_PATH
    cpus = host_cpus()
    cores = max(1, min(16, cpus["physical_cores"] or cpus["allowed"]))
    kind = "reference"
    kw = dict(p=PARAMS["p"], p2=PARAMS["p2"], lam=PARAMS["lam"], T0=PARAMS["T0"], Gamma=PARAMS["Gamma"])
    if os.path.exists(REF_FAST_PATH):
        os.environ.setdefault("OMP_PROC_BIND", "close")
        from oracle_lib import RefLib
        make = lambda grid, **k: RefLib(grid, path=REF_FAST_PATH, **k)  # noqa: E731
    else:
        from oracle_lib import Oracle
        make = Oracle
        kind, cores = "port", 1
    st = state

    def sample(threads, budget):
        if kind == "reference":
            _omp_threads(threads)
        cpu = make(g, **kw)
        cpu.set_history(st["t1grid"], st["dtratio"], st["rvec"], st["drvec"], st["QK"], st["QR"], st["dQK"],
                        st["dQR"])
        cpu.interpolate(0)
        cpu.set_control(control["delta_t"], control["delta"], control["delta_old"], int(control["loop"]),
                        int(control["method"]))
        n = 0
        t0 = time.perf_counter()
        while True:
            cpu.loop_iteration()
            n += 1
            el = time.perf_counter() - t0
            if el > budget or n >= 500:
                return n, el
    n1, el1 = sample(1, budget_s / 3)
    n, el = sample(cores, budget_s)
    log(f"[bench] CPU step at L=512: 1 thread {n1 / el1:.2f} steps/s, {cores} threads {n / el:.2f} steps/s")
    # the CPU-sized configuration: L=128 quench (grid Tmax 1e4), Gamma = 0.8, tmax = 1e3
    import dynamite_amd as dm
    g128 = dm.generate_grid(128, 1e4, 5, dm.INTERP_POLY, 9, 0)
    kw128 = dict(kw, Gamma=0.8)
    full = {}
    for threads in ((1, cores) if kind == "reference" else (1,)):
        if kind == "reference":
            _omp_threads(threads)
        cpu = make(g128, **kw128)
        cpu.init_quench()
        t0 = time.perf_counter()
        loops, t = 0, 0.0
        while t < 1e3:
            c = cpu.control()
            if c["delta_t"] < 1e-5:  # the main loop's delta_t >= delta_t_min condition
                break
            t = cpu.loop_iteration()
            loops += 1
        full[f"threads_{threads}_s"] = time.perf_counter() - t0
        full["loops"] = loops
    prm = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    ctx = new_context(g128, prm, rows_hint=4096)
    ctx.init_quench()
    t0 = time.perf_counter()
    steps, _ = ctx.run(1e3, 10**9)
    ctx.synchronize()
    full["gpu_dmfe_run_s"] = time.perf_counter() - t0
    full["gpu_loops"] = steps
    ctx.close()
    log(f"[bench] CPU-sized run (L=128 quench to its stop): {full}")
    return dict(value=n / el, unit="RK steps/s", cores=cores, kind=kind,
                sample=f"{n} main-loop RK steps of the same L=512 live state (T1={len(st['t1grid'])}) on "
                       f"{cores} OpenMP threads, {el:.1f}s wall; {n1} steps on 1 thread, {el1:.1f}s",
                threads_1_value=n1 / el1, host=cpus,
                cpu_sized_run=dict(workload="L=128 poly-9 grid (Tmax 1e4), p=3 q=4 lambda=0.5 Gamma=0.8 T0=inf, "
                                            "quench to tmax 1e3 (the main loop stops earlier by delta_t < "
                                            "delta_t_min)", **full))


def kernel_microbench(g, reps=200):
    """RHS and interpolation kernels on the seeded synthetic 512 x 512 history
    (QK, QR = exp(-|theta_k - theta_j| u_k) (1 + 1e-3 xi), u ~ U[0.5, 2], rows at
    t_k = 1e3 theta_k; dynamite_amd.workloads.synthetic_history)."""
    import dynamite_amd as dm
    from dynamite_amd.workloads import synthetic_history
    st = synthetic_history(g.theta)
    prm = dm.default_params(p=PARAMS["p"], p2=PARAMS["p2"], lambda_=PARAMS["lam"], T0=PARAMS["T0"],
                            Gamma=PARAMS["Gamma"])
    ctx = new_context(g, prm, rows_hint=len(st["t1grid"]) + 8)
    ctx.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
    ctx.interpolate(dm.POS_BSEARCH)
    ctx.synchronize()
    dm.profile(ctx, True)
    for _ in range(20):
        ctx.interpolate(dm.POS_ISEARCH)
    ctx.synchronize()
    s_ms, s_n = dm.profile_read(ctx, 0)
    dm.profile(ctx, False)
    rhs_ms = dm.time_rhs(ctx, reps)
    ctx.close()
    B = algorithmic_bytes_rhs(g.N, g.m)
    return dict(workload=f"synthetic {len(st['t1grid'])} x {g.N} history (seed {BATCH_SEED}), rows at "
                         f"t = 1e3 theta; L={g.N} poly-9 grid",
                rhs_us=rhs_ms * 1e3, rhs_achieved_GBs=B / (rhs_ms * 1e-3) / 1e9,
                rhs_frac=B / (rhs_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                interpolation_us=s_ms / max(s_n, 1) * 1e3)


def dense_microbench(N=512, order=9, B=4, reps=10):
    """Dense B-spline path: L=512 grid with interp-method bspline degree 9 (m = N), the
    A1 table applied to B=4 right-hand sides (QK and QR of the last two rows of the
    seeded synthetic history) as one (N^2 x N) . (N x B) FP64 MFMA contraction; plus one
    full RHS evaluation on the dense grid (A1/A2 on MFMA, B2 as the 64-lane stencil)."""
    import numpy as np
    import dynamite_amd as dm
    from dynamite_amd.workloads import synthetic_history
    t0 = time.time()
    g = dm.generate_grid(N, 1e5, 5, dm.INTERP_BSPLINE, order, 0, device=True)
    t_gen = time.time() - t0
    log(f"[bench] dense grid L={N} bspline {order} generated in {t_gen:.1f}s (weights on the device)")
    st = synthetic_history(g.theta)
    prm = dm.default_params(p=PARAMS["p"], p2=PARAMS["p2"], lambda_=PARAMS["lam"], T0=PARAMS["T0"],
                            Gamma=PARAMS["Gamma"])
    ctx = new_context(g, prm, rows_hint=len(st["t1grid"]) + 8)
    ctx.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
    ctx.interpolate(dm.POS_BSEARCH)
    QK, QR = st["QK"].reshape(-1, N), st["QR"].reshape(-1, N)
    X = np.stack([QK[-1], QR[-1], QK[-2], QR[-2]], axis=1)[:, :B]
    _, ms = dm.dense_apply(ctx, X, 0, reps)
    rhs_ms = dm.time_rhs(ctx, 5)
    ctx.close()
    P = N * N
    bytes_ = 8.0 * P * N + 8.0 * N * B + 8.0 * P * B
    flops = 2.0 * P * N * B
    return dict(workload=f"L={N} bspline degree {order} (m = N), A1 table x {B} right-hand sides "
                         f"(QK, QR of the last two synthetic rows)",
                contraction_us=ms * 1e3, bound="hbm", achieved_GBs=bytes_ / (ms * 1e-3) / 1e9,
                frac=bytes_ / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, achieved_TFLOPs=flops / (ms * 1e-3) / 1e12,
                fp64_mfma_peak_TFLOPs=78.6, dense_rhs_us=rhs_ms * 1e3, grid_generation_s=t_gen)


def sensitivity_run(tmax=1e4, h=1e-6, many_tmax=30.0):
    """Gradients of E and QK(t,0) at t_final w.r.t. (Gamma, lambda) for p=3+4, lambda=0.5,
    Gamma=0.7 on an L=256 poly-9 grid (forward-mode tangents carried with the run): the
    cost against the same run without tangents, and central finite differences (h = 1e-6)
    at the same t_final of runs replaying the nominal dt / integrator sequence (recorded
    from a host-driven nominal run, which makes the same steps as dmfe_run)."""
    import numpy as np
    import dynamite_amd as dm
    g = make_grid(256)

    def prm(G=0.7, lam=0.5):
        return dm.default_params(p=3, p2=4, lambda_=lam, T0=1e50, Gamma=G)
    c = new_context(g, prm(), rows_hint=1024)
    c.init_quench()
    t0 = time.perf_counter()
    n_fwd, _ = c.run(tmax, 10**12)
    c.synchronize()
    el_fwd = time.perf_counter() - t0
    c.close()
    c = new_context(g, prm(), rows_hint=1024)
    t0 = time.perf_counter()
    grad, t_final, n_grad = c.gradients(tmax, 10**12)
    c.synchronize()
    el_grad = time.perf_counter() - t0
    c.close()
    log(f"[bench] sensitivities: {n_grad} steps to t={t_final:.4g}: forward {el_fwd:.2f}s, with tangents "
        f"{el_grad:.2f}s ({el_grad / el_fwd:.2f}x); grad {grad}")
    # finite differences at t_final with the dt sequence frozen: the nominal sequence
    t0 = time.perf_counter()
    c = new_context(g, prm(), rows_hint=1024)
    c.init_quench()
    seq, t = [], 0.0
    while t < tmax:
        ctl = c.get_control()
        seq.append((ctl["delta_t"], ctl["method"]))
        t = c.loop_iteration()
    c.close()
    if len(seq) != n_grad or t != t_final:
        raise RuntimeError(f"host-driven nominal run: {len(seq)} steps to t={t!r}, dmfe_run {n_grad} to {t_final!r}")
    J = np.array([[grad[0], grad[1]], [grad[2], grad[3]]])

    def replay(G, lam):
        r = new_context(g, prm(G, lam), rows_hint=len(seq) + 64)
        r.init_quench()
        for dt, method in seq:
            r.set_control(dt, 1.0, 0.0, 10, method)
            r.rk_step()
        E, q0 = r.observables(t)
        r.close()
        return np.array([E, q0])
    Jfd = np.zeros((2, 2))
    for col, (dg, dl) in enumerate(((h, 0.0), (0.0, h))):
        Jfd[:, col] = (replay(0.7 + dg, 0.5 + dl) - replay(0.7 - dg, 0.5 - dl)) / (2 * h)
    err = float(np.max(np.abs(J - Jfd)) / np.max(np.abs(Jfd)))
    el_fd = time.perf_counter() - t0
    log(f"[bench] sensitivities vs finite differences at t={t:.4g} ({len(seq)} steps, {el_fd:.0f}s): "
        f"max rel err {err:.2e}")
    # 16 (Gamma, lambda) points (seeded), one after another and all at once
    rng = np.random.default_rng(BATCH_SEED + 16)
    pts = [(float(G), float(lam)) for G, lam in zip(rng.uniform(0.3, 1.0, 16), rng.uniform(0.0, 1.0, 16))]
    pl = [prm(G, lam) for G, lam in pts]
    t0 = time.perf_counter()
    seq_grads = []
    for p in pl:
        c = new_context(g, p, rows_hint=1024)
        seq_grads.append(c.gradients(many_tmax, 10**12)[0])
        c.close()
    el_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    grads_m, tf_m, st_m = dm.gradients_many([g] * len(pl), pl, many_tmax, 10**12)
    el_many = time.perf_counter() - t0
    seq_arr = np.array(seq_grads)
    same = bool(np.array_equal(seq_arr, grads_m))
    diff_pts = [i for i in range(len(pl)) if not np.array_equal(seq_arr[i], grads_m[i])]
    max_rel = float(np.max(np.abs(seq_arr - grads_m)) / np.max(np.abs(seq_arr)))
    if not same:
        log(f"[bench] 16 points: concurrent differs from sequential at points {diff_pts}, max rel diff {max_rel:.2e}")
    log(f"[bench] 16 points to t={many_tmax:g} ({int(st_m.sum())} steps): one after another {el_seq:.2f}s, "
        f"concurrent {el_many:.2f}s ({el_seq / el_many:.2f}x), identical={same}")
    # the 16 points as one batch: the tangents ride inside the batch's stage and RHS
    # launches (dmfe_batch_tangent_init), every instance with its own step control
    b = dm.Batch([g] * len(pl), pl, rows_cap=int(st_m.max()) + 64)
    t0 = time.perf_counter()
    b.init_quench()
    b.tangent_init([1.0, 0.0], [0.0, 1.0])
    b.set_limits(many_tmax, 10**12)
    nb = b.run(10**9, use_graph=True, chunk=16)
    bE, bQ = b.tangent_observables()
    el_batch = time.perf_counter() - t0
    b.close()
    grads_b = np.stack([bE[:, 0], bE[:, 1], bQ[:, 0], bQ[:, 1]], axis=1)
    same_b = bool(np.array_equal(seq_arr, grads_b))
    rel_b = float(np.max(np.abs(seq_arr - grads_b)) / np.max(np.abs(seq_arr)))
    log(f"[bench] 16 points in one batch: {el_batch:.2f}s ({nb} batch steps, {el_seq / el_batch:.2f}x one after "
        f"another), identical to the single-point gradients={same_b} (max rel diff {rel_b:.2e})")
    many = dict(points=len(pl), t_final=many_tmax, total_steps=int(st_m.sum()), sequential_s=el_seq,
                concurrent_s=el_many, speedup=el_seq / el_many, identical=same, max_rel_diff=max_rel,
                batch_s=el_batch, batch_steps=int(nb), batch_speedup=el_seq / el_batch, batch_identical=same_b,
                batch_max_rel_diff=rel_b,
                workload="Gamma ~ U[0.3,1], lambda ~ U[0,1] (seeded), same grid and model")
    return dict(workload=f"L=256 poly-9 grid (spline 5, Tmax 1e5), p=3+4, lambda=0.5, Gamma=0.7, quench to "
                         f"t_final={t_final:.4g} ({n_grad} adaptive RK steps)",
                grad={"dE_dGamma": grad[0], "dE_dlambda": grad[1], "dQK0_dGamma": grad[2], "dQK0_dlambda": grad[3]},
                mode="forward (tangent-linear), 2 directions, dt sequence frozen",
                forward_s=el_fwd, with_tangents_s=el_grad, cost_ratio=el_grad / el_fwd,
                fd_check={"t_final": t, "steps": len(seq), "h": h, "max_rel_err": err,
                          "jacobian_fd": Jfd.tolist(), "dt_sequence": "frozen (nominal run)"},
                batched=many)


BATCH_SEED = 20261203


def batch_workload(n=64):
    """64 quenches: bath temperature Gamma ~ U[0.3, 1.0], lambda ~ U[0, 1], L drawn 50/50
    from {256, 512} (seeded), p=3, q=4, T0=inf; every instance on its own grid package
    of the bench grid family (instances of one L share the device tables)."""
    import dynamite_amd as dm
    from dynamite_amd.workloads import batch_workload_spec
    Gam, lam, Ls = batch_workload_spec(n, BATCH_SEED)
    grids = {L: make_grid(int(L)) for L in sorted(set(Ls.tolist()))}
    gl, pl = [], []
    for i in range(n):
        gl.append(grids[int(Ls[i])])
        pl.append(dm.default_params(p=3, p2=4, lambda_=float(lam[i]), T0=1e50, Gamma=float(Gam[i])))
    return gl, pl, dict(Gamma=Gam, lam=lam, L=Ls)


def batch_run(prep, steps, looped_steps, chunk=16):
    """instance-steps/s of the batch (hipGraph replays) and of the plain path looped over
    the same instances (one context per instance, its steps one after another), both
    from the same live states after `prep` batch steps."""
    import numpy as np
    import dynamite_amd as dm
    gl, pl, meta = batch_workload()
    n = len(gl)
    b = dm.Batch(gl, pl, rows_cap=prep + steps + 64)
    b.init_quench()
    t0 = time.perf_counter()
    b.run(prep, use_graph=USE_GRAPHS, chunk=chunk)
    log(f"[bench] batch of {n}: {prep} prep steps in {time.perf_counter() - t0:.2f}s, "
        f"T1 {min(b.T1(i) for i in range(n))}..{max(b.T1(i) for i in range(n))}")
    # the batch autotuner picks the lane-group width of the batch RHS on these states
    G_b, table_b = b.autotune(reps=4)
    log(f"[bench] batch autotune: G={G_b}; ms per RHS launch {table_b}")
    b.run(2 * chunk, use_graph=USE_GRAPHS, chunk=chunk)  # warm-up, graph instantiation
    loops0 = np.array([b.get_control(i)["loop"] for i in range(n)])
    states = [(b.get_state(i), b.get_control(i)) for i in range(n)]
    t0 = time.perf_counter()
    nb = b.run(steps, use_graph=USE_GRAPHS, chunk=chunk)
    el = time.perf_counter() - t0
    loops1 = np.array([b.get_control(i)["loop"] for i in range(n)])
    inst_steps = int((loops1 - loops0).sum())
    b.close()
    batch_rate = inst_steps / el
    log(f"[bench] batch: {nb} batch steps, {inst_steps} instance-steps in {el:.3f}s -> {batch_rate:.0f}/s "
        f"({1e3 * el / max(nb, 1):.2f} ms per batch step)")
    # looped plain path on the same states, one context per instance advanced on its own
    # (one instance per launch): with dmfe_run (device-side step control, the faster
    # plain path) and with the host-driven dmfe_loop_iteration; the first context of each
    # grid size is autotuned and installs its shape for the contexts created after it
    ctxs, tuned = [], {}
    for i in range(n):
        st, c = states[i]
        ctx = new_context(gl[i], pl[i], rows_hint=len(st["t1grid"]) + 3 * looped_steps + 64)
        ctx.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
        ctx.set_control(c["delta_t"], c["delta"], c["delta_old"], c["loop"], c["method"])
        ctx.interpolate(dm.POS_BSEARCH)
        if gl[i].N not in tuned:
            tuned[gl[i].N] = dm.autotune(ctx, reps=10)[0]
        ctx.run(1e300, c["loop"] + looped_steps)  # warm-up, graph capture
        ctx.synchronize()
        ctxs.append(ctx)
    log(f"[bench] looped path shapes (G, split) by L: {tuned}")
    t0 = time.perf_counter()
    for ctx in ctxs:
        ctx.run(1e300, ctx.get_control()["loop"] + looped_steps)
        ctx.synchronize()
    el_l = time.perf_counter() - t0
    t0 = time.perf_counter()
    for ctx in ctxs:
        for _ in range(looped_steps):
            ctx.loop_iteration()
        ctx.synchronize()
    el_h = time.perf_counter() - t0
    for ctx in ctxs:
        ctx.close()
    looped_rate = n * looped_steps / el_l
    host_rate = n * looped_steps / el_h
    log(f"[bench] looped plain path: {n} x {looped_steps} steps: dmfe_run {el_l:.3f}s -> {looped_rate:.0f}/s, "
        f"host-driven loop {el_h:.3f}s -> {host_rate:.0f}/s")
    best = max(looped_rate, host_rate)
    return dict(instances=n, value=batch_rate, unit="instance-steps/s", looped_value=best,
                ratio=batch_rate / best, looped_dmfe_run_value=looped_rate, looped_host_driven_value=host_rate,
                looped_steps=looped_steps, batch_steps=nb, instance_steps=inst_steps,
                ms_per_batch_step=1e3 * el / max(nb, 1), prep_steps=prep, graph_chunk=chunk,
                autotune={"batch_G": G_b, "batch_rhs_ms": {f"G{G}": round(ms, 4) for G, ms in table_b.items()},
                          "looped_shapes": {f"L{L}": {"G": sh[0], "split": sh[1]} for L, sh in tuned.items()}},
                workload=f"{n} quenches, Gamma~U[0.3,1], lambda~U[0,1], L 50/50 in {{256,512}} "
                         f"(seed {BATCH_SEED}), p=3 q=4 T0=inf, poly-9 grids (spline 5, Tmax 1e5); "
                         f"timed from the live states after {prep} batch steps")


def read_traffic():
    """HBM bytes per rhs_rows_kernel launch on this workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/rhs_rows_kernel_L512_hbm.txt: the line
    `traffic_bytes_per_launch = ...`, counters scaled by the plain-copy calibration)."""
    path = os.path.join(ROOT, "profiles", "rhs_rows_kernel_L512_hbm.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1].split()[0])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--prep", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-dense", action="store_true")
    ap.add_argument("--no-sensitivity", action="store_true")
    ap.add_argument("--batch-prep", type=int, default=1000)
    ap.add_argument("--batch-steps", type=int, default=64)
    ap.add_argument("--looped-steps", type=int, default=48)
    ap.add_argument("--batch-only", action="store_true", help="only the batch-64 measurement")
    ap.add_argument("--sensitivity-only", action="store_true", help="only the sensitivity measurement")
    ap.add_argument("--no-graphs", action="store_true",
                    help="issue the step launches directly instead of replaying hipGraphs (same kernels)")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a plain HBM copy of this many MiB (counter calibration) and exit")
    ap.add_argument("--rhs-only", type=int, default=0,
                    help="only prepare the live L=512 state (--prep steps) and launch this many RHS "
                         "evaluations back to back on it (for counter runs), then exit")
    ap.add_argument("--dense-only", action="store_true", help="only the dense B-spline measurement")
    ap.add_argument("--full-run", type=float, default=0.0,
                    help="only integrate the L=512 configuration from the quench to this t (1e5 for the "
                         "production length) with dmfe_run, reporting every 10000 loops, then exit")
    a = ap.parse_args()
    global USE_GRAPHS
    USE_GRAPHS = not a.no_graphs
    if a.copy_probe > 0:
        import dynamite_amd as dm
        nb = int(a.copy_probe * 2**20)
        ms = dm.copy_probe(nb, 10)
        log(f"[bench] copy probe: {nb} bytes read + {nb} written, {ms * 1e3:.1f} us/launch, "
            f"{2 * nb / ms / 1e6:.0f} GB/s")
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {a.gpus}); independent "
              "quenches would shard across GPUs with no communication, which is not run here.",
              file=sys.stderr)
        sys.exit(2)
    if a.sensitivity_only:
        print(json.dumps({"sensitivities": sensitivity_run()}), flush=True)
        return
    if a.dense_only:
        print(json.dumps({"dense_bspline": dense_microbench()}), flush=True)
        return
    if a.batch_only:
        out = batch_run(a.batch_prep, a.batch_steps, a.looped_steps)
        print(json.dumps({"batch": out}), flush=True)
        return
    N = 512
    g = make_grid(N)
    if a.rhs_only > 0:
        print(json.dumps({"rhs_only": rhs_only_run(g, a.prep, a.rhs_only)}), flush=True)
        return
    if a.full_run > 0:
        print(json.dumps({"full_run": full_run(g, a.full_run)}), flush=True)
        return
    log(f"[bench] GPU: {a.prep} prep + {a.warmup} warmup + {a.steps} timed steps")
    el, prof, state, control, T1, rhs_b2b_ms, tune, host = gpu_run(g, a.prep, a.steps, a.warmup)
    ms_step = 1e3 * el / a.steps
    rhs_ms, rhs_n = prof[1]
    avg_rhs_in_steps_s = rhs_ms / 1e3 / max(rhs_n, 1)
    avg_rhs_s = rhs_b2b_ms / 1e3
    B = algorithmic_bytes_rhs(N, g.m)
    achieved = B / avg_rhs_s / 1e9
    log(f"[bench] {ms_step:.3f} ms/step; rhs_rows {rhs_n} launches avg {avg_rhs_in_steps_s * 1e6:.1f} us in steps, "
        f"{avg_rhs_s * 1e6:.1f} us back to back "
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
        # no dataset: the state is the quench's own initial condition integrated --prep steps
        "data": "synthetic (quench from its initial condition, then the timed steps on that live state)",
        "config": {"workload": "plain quench, L=512 poly-9 grid (spline 5, Tmax 1e5), p=3 q=4 lambda=0.5 "
                               "Gamma=0.6 T0=inf; adaptive RK steps (dmfe_run) on the live state after the prep steps",
                   "L": N, "stencil_width": g.m, "prep_steps": a.prep, "T1_end": T1,
                   "method": "RK54" if control["method"] == 1 else "SSPRK104", "delta_t": control["delta_t"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "rhs_rows_kernel", "bytes_per_launch": B,
                     "avg_launch_us": avg_rhs_s * 1e6, "avg_launch_us_in_steps": avg_rhs_in_steps_s * 1e6,
                     "launches_in_steps": rhs_n},
        "autotune": tune,
        "host_driven_loop": host,
    }
    out["kernel_microbench"] = kernel_microbench(g)
    log(f"[bench] microbench: {out['kernel_microbench']}")
    if not a.no_dense:
        out["dense_bspline"] = dense_microbench()
        log(f"[bench] dense: {out['dense_bspline']}")
    if not a.no_sensitivity:  # before the batch, whose autotuning installs L=256 shapes
        out["sensitivities"] = sensitivity_run()
    if not a.no_batch:
        out["batch"] = batch_run(a.batch_prep, a.batch_steps, a.looped_steps)
    if not a.no_cpu:
        log(f"[bench] CPU baseline (budget {a.cpu_budget:.0f}s)")
        out["cpu_baseline"] = cpu_baseline(g, state, control, a.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
