"""The C oracle's main loop against the reference's own: the L=128 quench (p=3, p2=4,
lambda=0.5, Gamma=0.8, T0=inf; poly-9 grid, spline 5, Tmax 1e4) with tmax = 1e3.

The golden fixture tests/golden/quench128_main_loop.npz was written by
tests/golden/make_quench128_fixture.py from the reference's CPU sources
(oracle/_ref/libdmfe_ref_strict.so, one OpenMP thread): (loop, t, E(t), QK(t,0)) every
25 iterations and at the last.  The loop (simulation_runner.cpp:125-304) ends by its
delta_t >= delta_t_min condition at loop 2061, t = 4.037, after the step has shrunk by
0.9 for every error above 2 delta_max; the oracle must reproduce every record bit for
bit, the stop included.  About 30 s of CPU.
"""
import os

import numpy as np

import dynamite_amd as dm
from oracle_lib import Oracle

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "quench128_main_loop.npz")


def test_oracle_main_loop_matches_reference_fixture():
    f = np.load(FIXTURE)
    N, tmax_grid, spline, order, p, p2, lam, gamma, T0, dt_min, delta_max, tmax = f["params"]
    N = int(N)
    g = dm.generate_grid(N, tmax_grid, int(spline), dm.INTERP_POLY, int(order), 0)
    o = Oracle(g, p=int(p), p2=int(p2), lam=lam, T0=T0, Gamma=gamma, dt_min=dt_min, delta_max=delta_max)
    o.init_quench()
    want = {int(l): i for i, l in enumerate(f["loop"])}
    loop, running = 0, True
    while running:
        t = o.loop_iteration()
        loop += 1
        running = t < tmax and o.control()["delta_t"] >= dt_min
        if loop in want:
            i = want[loop]
            assert t == f["t"][i], (loop, t, f["t"][i])
            assert o.energy(t) == f["energy"][i], loop
            assert o.state()["QK"][-N] == f["qk0"][i], loop
    c = o.control()
    assert loop == int(f["final_control"][3]) == int(f["loop"][-1])
    assert c["delta_t"] == f["final_control"][0] and c["method"] == int(f["final_control"][4])
    o.close()


def test_reference_thread_count_sensitivity_near_the_stop():
    """The same reference run on 1 and on 8 OpenMP threads (quench128_main_loop.npz and
    quench128_main_loop_8threads.npz, both from tests/golden/make_quench128_fixture.py):
    the loop count and every proposed time agree, E and QK(t,0) agree to rounding while
    the step is well inside its stability region (loop <= 1375, t <= 1.6) and differ by
    ~1e-6 near the stop, where the step sits at its stability limit.  This is the spread
    of the reference's own result under another summation order, which bounds how close
    any other implementation can be expected to follow it there."""
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    a = np.load(os.path.join(d, "quench128_main_loop.npz"))
    b = np.load(os.path.join(d, "quench128_main_loop_8threads.npz"))
    assert np.array_equal(a["loop"], b["loop"]) and np.array_equal(a["t"], b["t"])
    # the final step size, loop count and integrator agree (the error norms differ in their last bits)
    assert np.array_equal(a["final_control"][[0, 3, 4]], b["final_control"][[0, 3, 4]])
    early = a["loop"] <= 1375
    for key in ("energy", "qk0"):
        err = np.abs(a[key] - b[key]) / np.abs(a[key])
        assert err[early].max() < 1e-13, key
        assert 1e-7 < err[~early].max() < 1e-5, (key, err[~early].max())
