"""Writes tests/golden/maint256_main_loop.npz: the reference's own CPU main loop on the
L=256 quench (p=3, p2=4, lambda=0.5, Gamma=0.7, T0=inf; poly-9 grid, spline order 5,
Tmax=1e5; delta_max=1e-10, delta_t_min=1e-5) from the initial condition through loop
100000 - where the main loop runs its maintenance (sparsifyNscale with threshold
delta_max, interpolate, and the dt halving with the switch to SSPRK104 when
delta < delta_max / 2; simulation_runner.cpp:161-233, sparsify_utils.cpp:14-97) - and
RECORD_AFTER further loops.  Recorded: (loop, t, E(t), QK(t,0)) every EVERY loops and at
every loop from 99990 on; T1, delta_t, delta and the integrator just before and just
after the maintenance; the t1grid kept by the maintenance.

The reference runs from its sources compiled by oracle/Makefile into
oracle/_ref/libdmfe_ref_fast.so (the reference's CPU flags, -O3 -ffast-math, OpenMP) on
THREADS threads: the strict one-thread build would take hours for 1e5 loops at L=256.
OpenMP reductions combine partial sums in thread arrival order, so the records carry
rounding noise of the reference itself at the 1e-13 level.  Run from the repository
root after `make -C oracle ref`:

    python3 tests/golden/make_maint256_fixture.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

N, TMAX_GRID, SPLINE, ORDER = 256, 1e5, 5, 9
P, P2, LAM, GAMMA, T0 = 3, 4, 0.5, 0.7, 1e50
DT_MIN, DELTA_MAX = 1e-5, 1e-10
MAINT = 100000
RECORD_AFTER = 200
EVERY = 1000
THREADS = 7
OUT = os.path.join(HERE, "maint256_main_loop.npz")


def main():
    os.environ["OMP_NUM_THREADS"] = str(THREADS)
    import dynamite_amd as dm
    from oracle_lib import REF_FAST_PATH, RefLib

    if not os.path.exists(REF_FAST_PATH):
        raise SystemExit("oracle/_ref/libdmfe_ref_fast.so is missing: run `make -C oracle ref` first")
    g = dm.generate_grid(N, TMAX_GRID, SPLINE, dm.INTERP_POLY, ORDER, 0)
    ref = RefLib(g, p=P, p2=P2, lam=LAM, T0=T0, Gamma=GAMMA, dt_min=DT_MIN, delta_max=DELTA_MAX, path=REF_FAST_PATH)
    ref.init_quench()
    rows, before, after, kept = [], None, None, None
    t0 = time.time()
    loop = 0
    while loop < MAINT + RECORD_AFTER:
        if loop == MAINT - 1:
            c = ref.control()
            before = (ref.T1 + 1, c["delta_t"], c["delta"], c["method"])  # T1 after the next append
        t = ref.loop_iteration()
        loop += 1
        if loop == MAINT:
            c = ref.control()
            after = (ref.T1, c["delta_t"], c["delta"], c["method"])
            kept = ref.t1grid()
        if loop % EVERY == 0 or loop >= MAINT - 10:
            rows.append((loop, t, ref.energy(t), ref.last_qk0()))
        if loop % 5000 == 0:
            print(f"loop {loop} t {t:.6g} dt {ref.control()['delta_t']:.4g} T1 {ref.T1} ({time.time() - t0:.0f} s)",
                  flush=True)
        if not ref.control()["delta_t"] >= DT_MIN:
            raise SystemExit(f"the main loop stopped by delta_t < delta_t_min at loop {loop}")
    rec = np.array(rows)
    np.savez(OUT, loop=rec[:, 0].astype(np.int64), t=rec[:, 1], energy=rec[:, 2], qk0=rec[:, 3],
             before=np.array(before), after=np.array(after), kept_t1grid=kept,
             params=np.array([N, TMAX_GRID, SPLINE, ORDER, P, P2, LAM, GAMMA, T0, DT_MIN, DELTA_MAX, MAINT, THREADS]))
    print(f"{loop} iterations; before maintenance T1/dt/delta/method {before}, after {after}; "
          f"{len(rows)} records -> {OUT}", flush=True)


if __name__ == "__main__":
    main()
