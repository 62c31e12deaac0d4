"""Writes tests/golden/ref_gradient256.npz: the gradient of E(t_f) and QK(t_f, 0) with
respect to (Gamma, lambda) computed by the reference's own CPU path, for the sensitivity
workload (p = 3 + 4, lambda = 0.5, Gamma = 0.7, T0 = inf) on the L = 256 poly-9 grid
(spline 5, Tmax 1e5) from the quench to t_f ~ 100.

The reference sources compiled by oracle/Makefile (oracle/_ref/libdmfe_ref_strict.so,
one OpenMP thread) run the main loop (simulation_runner.cpp:125-304) once with the
nominal parameters, which fixes the step sequence (delta_t and integrator of every
step); four more runs replay that sequence with Gamma +- h and lambda +- h (h = 1e-6)
through update() (runge_kutta.cpp:254-261), and the central differences of the energy
line (simulation_runner.cpp:334-337) and of QK(t_f, 0) give the Jacobian.  The fixture
holds the Jacobian, the step sequence and the nominal observables.

    python tests/golden/make_ref_gradient256.py      (needs oracle/_ref, a few minutes)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
os.environ["OMP_NUM_THREADS"] = "1"

import dynamite_amd as dm  # noqa: E402  (grid generator, host code only)
from oracle_lib import RefLib  # noqa: E402

N, TMAX_GRID, SPLINE, ORDER = 256, 1e5, 5, 9
P, P2, LAM, GAMMA, T0 = 3, 4, 0.5, 0.7, 1e50
T_FINAL, H = 100.0, 1e-6


def main():
    g = dm.generate_grid(N, TMAX_GRID, SPLINE, dm.INTERP_POLY, ORDER, 0)
    ref = RefLib(g, p=P, p2=P2, lam=LAM, T0=T0, Gamma=GAMMA)
    ref.init_quench()
    seq, t = [], 0.0
    t0 = time.time()
    while t < T_FINAL:
        c = ref.control()
        seq.append((c["delta_t"], c["method"]))
        t = ref.loop_iteration()
    st = ref.state()
    E0, Q0 = ref.energy(t), st["QK"][(len(st["t1grid"]) - 1) * N]
    print(f"nominal: {len(seq)} steps to t = {t!r} in {time.time() - t0:.0f} s, E = {E0!r}, QK0 = {Q0!r}", flush=True)

    def replay(G, lam):
        r = RefLib(g, p=P, p2=P2, lam=lam, T0=T0, Gamma=G)
        r.init_quench()
        for dt, method in seq:
            r.set_control(dt, 1.0, 0.0, 10, method)
            r.update()
        s = r.state()
        return np.array([r.energy(t), s["QK"][(len(s["t1grid"]) - 1) * N]])

    J = np.zeros((2, 2))
    for col, (dG, dL) in enumerate(((H, 0.0), (0.0, H))):
        J[:, col] = (replay(GAMMA + dG, LAM + dL) - replay(GAMMA - dG, LAM - dL)) / (2 * H)
        print(f"column {col} done ({time.time() - t0:.0f} s)", flush=True)
    seq = np.array(seq)
    np.savez(os.path.join(HERE, "ref_gradient256.npz"),
             params=np.array([N, TMAX_GRID, SPLINE, ORDER, P, P2, LAM, GAMMA, T0, H]),
             t_final=np.array(t), dt=seq[:, 0], method=seq[:, 1].astype(np.int32),
             observables=np.array([E0, Q0]), jacobian=J)
    print("J = [[dE/dGamma, dE/dlambda], [dQK0/dGamma, dQK0/dlambda]] =", J.tolist(), flush=True)


if __name__ == "__main__":
    main()
