"""Writes tests/golden/quench128_main_loop.npz: the reference's own CPU main loop on the
L=128 quench with tmax = 1e3 (p=3, p2=4, lambda=0.5, Gamma=0.8, T0=inf; poly-9 grid,
spline order 5, Tmax=1e4; delta_max=1e-10, delta_t_min=1e-5), recording (loop, t,
E(t), QK(t,0)) every 25 iterations and at the last one.  The loop runs while
t < tmax and delta_t >= delta_t_min (simulation_runner.cpp:125); on this configuration
the error stays near delta_max, the step shrinks by 0.9 whenever delta > 2 delta_max, and
the loop ends by the delta_t_min condition at loop 2061, t = 4.037.

The reference runs from its sources compiled by oracle/Makefile into
oracle/_ref/libdmfe_ref_strict.so (the loop of simulation_runner.cpp:125-304 through
oracle/ref_harness.cpp, observables as in simulation_runner.cpp:327-341).  Run from the
repository root after `make -C oracle ref`:

    python3 tests/golden/make_quench128_fixture.py            # one OpenMP thread
    python3 tests/golden/make_quench128_fixture.py 8          # the same run on 8 threads

The 8-thread run (quench128_main_loop_8threads.npz) is the same reference build with its
OpenMP reductions split over 8 threads: it measures how far the reference's own result
moves under another summation order.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

N, TMAX_GRID, SPLINE, ORDER = 128, 1e4, 5, 9
P, P2, LAM, GAMMA, T0 = 3, 4, 0.5, 0.8, 1e50
DT_MIN, DELTA_MAX = 1e-5, 1e-10
T_FINAL = 1e3
EVERY = 25
THREADS = int(sys.argv[1]) if len(sys.argv) > 1 else 1
OUT = os.path.join(HERE, "quench128_main_loop.npz" if THREADS == 1 else f"quench128_main_loop_{THREADS}threads.npz")


def main():
    # one thread by default: the reference's OpenMP reductions combine partial sums in
    # thread arrival order, so more threads make the last bits (and, near the stall, the
    # trajectory) vary
    os.environ["OMP_NUM_THREADS"] = str(THREADS)
    import dynamite_amd as dm
    from oracle_lib import RefLib, ref_available

    if not ref_available():
        raise SystemExit("oracle/_ref/libdmfe_ref_strict.so is missing: run `make -C oracle ref` first")
    g = dm.generate_grid(N, TMAX_GRID, SPLINE, dm.INTERP_POLY, ORDER, 0)
    ref = RefLib(g, p=P, p2=P2, lam=LAM, T0=T0, Gamma=GAMMA, dt_min=DT_MIN, delta_max=DELTA_MAX)
    ref.init_quench()
    rows = []
    t_last, loop = 0.0, 0
    t0 = time.time()
    running = True
    while running:
        t = ref.loop_iteration()
        loop += 1
        t_last = t  # the step appended the row at the proposed time
        running = t_last < T_FINAL and ref.control()["delta_t"] >= DT_MIN
        if loop % EVERY == 0 or not running:
            qk0 = ref.state()["QK"][-N]  # row-major T1 x N: column 0 of the last row
            rows.append((loop, t, ref.energy(t), qk0))
        if loop % 1000 == 0:
            print(f"loop {loop} t {t:.6g} ({time.time() - t0:.0f} s)", flush=True)
    rec = np.array(rows)
    c = ref.control()
    np.savez(OUT, loop=rec[:, 0].astype(np.int64), t=rec[:, 1], energy=rec[:, 2], qk0=rec[:, 3],
             final_control=np.array([c["delta_t"], c["delta"], c["delta_old"], c["loop"], c["method"]]),
             params=np.array([N, TMAX_GRID, SPLINE, ORDER, P, P2, LAM, GAMMA, T0, DT_MIN, DELTA_MAX, T_FINAL]))
    print(f"{loop} iterations to t = {t_last:.17g}; E = {rec[-1, 2]:.17g}, QK(t,0) = {rec[-1, 3]:.17g}; "
          f"{len(rows)} records -> {OUT}", flush=True)


if __name__ == "__main__":
    main()
