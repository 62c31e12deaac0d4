"""Batched quenches (dmfe_batch_*) against independent single-instance contexts.

A batch runs the same kernels as a context (the stage, RHS and finish device code is
shared) with the scalar step control moved onto the device, so every instance of a
batch must reproduce dmfe_run of a context with its own parameters bit for bit:
identical time grids, histories, control scalars and correlation records.  Instances
differ in grid (two sizes), temperature T0, lambda and Gamma; the maintenance cadence
(sparsify + re-interpolation + dt halving, simulation_runner.cpp:161-278) is lowered so
that it runs inside the tested window: at loop 40 the halved dt falls below
delta_t_min and the main-loop condition stops the instance; the instances with the
loop-150 maintenance start from dt = 4 delta_t_min (set_control) and continue with
SSPRK104 after it.  The history arena starts small so that it grows.
"""
import numpy as np
import pytest

import dynamite_amd as dm

pytestmark = pytest.mark.gpu

_GRIDS = {}


def grid(N):
    if N not in _GRIDS:
        _GRIDS[N] = dm.generate_grid(N, 1e2, 5, dm.INTERP_POLY, 9, 0)
    return _GRIDS[N]


def instances():
    """(grid, params) of six quenches: two grid sizes, several temperatures and mixtures."""
    specs = [(64, 0.5, 1e50, 0.8, 150), (64, 0.3, 1.0, 0.5, 40), (96, 0.7, 0.6, 0.0, 150),
             (64, 0.9, 1e50, 0.2, 150), (96, 0.1, 2.0, 0.9, 40), (64, 0.5, 0.8, 0.3, 150)]
    out = []
    for N, lam, T0, G, every in specs:
        p = dm.default_params(p=3, p2=4, lambda_=lam, T0=T0, Gamma=G)
        p.maint_every = every
        p.delta_max = 1e-8
        out.append((grid(N), p))
    return out


DT0 = 4e-5


def context_run(g, p, steps):
    ctx = dm.Context(g, p, rows_hint=64)
    ctx.init_quench()
    if p.maint_every > 100:
        c = ctx.get_control()
        ctx.set_control(DT0, c["delta"], c["delta_old"], c["loop"], c["method"])
    n, corr = ctx.run(float("inf"), steps, corr_cap=steps)
    st, ctl = ctx.get_state(), ctx.get_control()
    ctx.close()
    return st, ctl, corr


def assert_same_state(sb, sc, what):
    for k in ("t1grid", "dtratio", "rvec", "drvec", "QK", "QR", "dQK", "dQR"):
        assert sb[k].shape == sc[k].shape, f"{what} {k}: {sb[k].shape} vs {sc[k].shape}"
        if not np.array_equal(sb[k], sc[k]):
            d = np.abs(sb[k] - sc[k]).max()
            raise AssertionError(f"{what} {k}: max |diff| {d:.3e}")


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_contexts(use_graph):
    steps = 170  # crosses the maintenance points and grows the arena past 64 and 128 rows
    inst = instances()
    print(f"batch of {len(inst)}, {steps} steps, graph={use_graph}", flush=True)
    b = dm.Batch([g for g, _ in inst], [p for _, p in inst], rows_cap=64)
    b.init_quench()
    for i, (g, p) in enumerate(inst):
        if p.maint_every > 100:
            c = b.get_control(i)
            b.set_control(i, DT0, c["delta"], c["delta_old"], c["loop"], c["method"])
    b.set_limits(float("inf"), steps)
    b.record(256)
    if use_graph:
        n = b.run(10 * steps, use_graph=True, chunk=8)
    else:
        n = 0
        while b.step() > 0:
            n += 1
    assert n >= steps
    stopped = 0
    for i, (g, p) in enumerate(inst):
        sc, cc, corr = context_run(g, p, steps)
        sb, cb = b.get_state(i), b.get_control(i)
        assert cb["status"] == dm.BATCH_DONE
        for k in ("delta_t", "delta", "delta_old", "loop", "method"):
            assert cb[k] == cc[k], (i, k, cb[k], cc[k])
        assert_same_state(sb, sc, f"instance {i}")
        stopped += cc["loop"] < steps
        rec = b.records(i)
        for lp in range(1, cc["loop"] + 1):
            row = rec[(lp - 1) % 256]
            assert row[3] == lp
            assert row[0] == corr[lp - 1, 0] and row[2] == corr[lp - 1, 2], (i, lp)
            assert row[1] == corr[lp - 1, 1], (i, lp, row[1], corr[lp - 1, 1])
    assert stopped == 2, stopped  # the two instances with the loop-40 maintenance
    b.close()


def test_graph_replay_matches_direct_launches():
    """Replays of a captured multi-step graph, with host synchronisation and the
    maintenance between replays, reproduce direct launches."""
    inst = instances()[:4]
    out = []
    for use_graph in (False, True):
        b = dm.Batch([g for g, _ in inst], [p for _, p in inst], rows_cap=64)
        b.init_quench()
        b.set_limits(float("inf"), 60)
        b.run(60, use_graph=use_graph, chunk=7)
        out.append([b.get_state(i) for i in range(len(inst))])
        b.close()
    for i in range(len(inst)):
        assert_same_state(out[1][i], out[0][i], f"instance {i}")


def test_batch_limits_stop_each_instance():
    inst = [instances()[k] for k in (0, 2, 3)]
    b = dm.Batch([g for g, _ in inst], [p for _, p in inst], rows_cap=64)
    b.init_quench()
    b.set_limits(float("inf"), 12)
    n = b.run(100, use_graph=True, chunk=5)
    assert n >= 12
    for i in range(3):
        c = b.get_control(i)
        assert c["loop"] == 12 and c["status"] == dm.BATCH_DONE
    # raising the limit resumes from where the instances stopped
    b.set_limits(float("inf"), 20)
    b.run(100, use_graph=False)
    assert all(b.get_control(i)["loop"] == 20 for i in range(3))
    # a time limit stops an instance once its proposed time reaches tmax
    tmax = 1.5 * max(b.get_state(i)["t1grid"][-1] for i in range(3))
    b.set_limits(tmax, 10**9)
    b.run(100000, use_graph=True, chunk=16)
    for i in range(3):
        c = b.get_control(i)
        assert c["status"] == dm.BATCH_DONE
        t = b.get_state(i)["t1grid"]
        # the last step was proposed below tmax and ended at or beyond it
        assert t[-2] < tmax and t[-1] >= tmax * (1 - 1e-12), (i, t[-2], t[-1], tmax)
    b.close()
