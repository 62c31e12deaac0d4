"""dmfe_run on the device-side step control against the host-driven loop.

dmfe_run advances a context with the dt rule, the loop condition and the stage
arguments evaluated on the device (hipGraph replays of 16 steps; the host services arena
growth and the maintenance cadence between replays), while dmfe_loop_iteration keeps
the host-driven step of the reference's main loop (simulation_runner.cpp:125-341).
Both must give bit-identical histories, control scalars and correlation rows, across
arena growth, the sparsify + dt-halving maintenance (cadence lowered to reach it) and
the SSPRK104 -> RK54 switch, from a quench and from set_state (no interpolation yet).
"""
import numpy as np
import pytest

import dynamite_amd as dm

pytestmark = pytest.mark.gpu

_GRIDS = {}


def grid(N, Tmax=1e2):
    if (N, Tmax) not in _GRIDS:
        _GRIDS[(N, Tmax)] = dm.generate_grid(N, Tmax, 5, dm.INTERP_POLY, 9, 0)
    return _GRIDS[(N, Tmax)]


def params(maint=0, Gamma=0.8, lam=0.5, T0=1e50):
    p = dm.default_params(p=3, p2=4, lambda_=lam, T0=T0, Gamma=Gamma)
    if maint:
        p.maint_every = maint
        p.delta_max = 1e-8
    return p


def host_run(ctx, steps):
    rows = []
    for _ in range(steps):
        t = ctx.loop_iteration()
        E, q0 = ctx.observables(t)
        rows.append((t, E, q0))
    return np.array(rows)


def same_state(a, b):
    sa, sb = a.get_state(), b.get_state()
    for k in ("t1grid", "dtratio", "rvec", "drvec", "QK", "QR", "dQK", "dQR"):
        assert sa[k].shape == sb[k].shape and np.array_equal(sa[k], sb[k]), k
    ca, cb = a.get_control(), b.get_control()
    for k in ("delta_t", "delta", "delta_old", "loop", "method"):
        assert ca[k] == cb[k], (k, ca[k], cb[k])


@pytest.mark.parametrize("N,maint,steps", [(64, 0, 240), (96, 150, 260)])
def test_run_matches_host_loop_from_quench(N, maint, steps):
    print(f"device loop vs host loop: N={N} maint={maint} steps={steps}", flush=True)
    g = grid(N)
    dev = dm.Context(g, params(maint), rows_hint=64)  # small arena: grows during the run
    host = dm.Context(g, params(maint), rows_hint=64)
    for c in (dev, host):
        c.init_quench()
        if maint:  # start above dt_min so that the dt halving at the maintenance keeps it running
            k = c.get_control()
            c.set_control(4e-5, k["delta"], k["delta_old"], k["loop"], k["method"])
    n, corr = dev.run(float("inf"), steps, corr_cap=steps)
    ref = host_run(host, n)
    assert n == steps
    assert np.array_equal(corr, ref)
    same_state(dev, host)
    if maint:
        assert dev.get_control()["method"] in (dm.RK54, dm.SSPRK104)
    dev.close()
    host.close()


def test_run_in_pieces_and_from_set_state():
    g = grid(64)
    a = dm.Context(g, params(), rows_hint=64)
    a.init_quench()
    a.run(float("inf"), 100)
    st, ctl = a.get_state(), a.get_control()
    n1, c1 = a.run(float("inf"), 180, corr_cap=200)
    b = dm.Context(g, params(), rows_hint=32)
    b.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
    b.set_control(ctl["delta_t"], ctl["delta"], ctl["delta_old"], ctl["loop"], ctl["method"])
    ref = host_run(b, n1)
    assert n1 == 80 and np.array_equal(c1, ref)
    same_state(a, b)
    # limits: tmax stops at the first proposed time >= tmax, as the host loop does
    t_stop = float(a.get_state()["t1grid"][-1]) * 1.5
    n2, _ = a.run(t_stop, 10**9)
    t = 0.0
    k = 0
    while True:
        t = b.loop_iteration()
        k += 1
        if not t < t_stop:
            break
    assert n2 == k
    same_state(a, b)
    a.close()
    b.close()


def test_run_with_tangents_matches_host_loop():
    """Tangents on the device loop (tangent RHS and finish launches in the replayed
    step), in segments across arena growth: primal and tangent histories and the tangent
    observables equal the host-driven loop's bit for bit."""
    g = grid(64)
    dev = dm.Context(g, params(Gamma=0.7), rows_hint=32)
    host = dm.Context(g, params(Gamma=0.7), rows_hint=32)
    for c in (dev, host):
        c.init_quench()
        c.tangent_init([1.0, 0.3], [0.0, -2.0])
    n, corr = dev.run(float("inf"), 300, corr_cap=300)
    ref = host_run(host, n)
    assert n == 300 and np.array_equal(corr, ref)
    same_state(dev, host)
    for k in range(2):
        a, b = dev.tangent_state(k), host.tangent_state(k)
        for key in ("QK", "QR", "rvec"):
            assert np.array_equal(a[key], b[key]), (k, key)
    t = corr[-1, 0]
    for x, y in zip(dev.tangent_observables(t), host.tangent_observables(t)):
        assert np.array_equal(x, y)
    dev.close()
    host.close()


def test_run_with_tangents_stops_before_sparsification():
    g = grid(64)
    out = []
    for use_run in (True, False):
        c = dm.Context(g, params(maint=50, Gamma=0.7), rows_hint=64)
        c.init_quench()
        k = c.get_control()
        c.set_control(4e-5, k["delta"], k["delta_old"], k["loop"], k["method"])
        c.tangent_init([1.0], [0.0])
        with pytest.raises(dm.DmfeError, match="sparsification point"):
            if use_run:
                c.run(float("inf"), 200)
            else:
                for _ in range(200):
                    c.loop_iteration()
        assert c.get_control()["loop"] == 49
        out.append((c.get_state(), c.tangent_state(0)))
        c.close()
    for k in ("t1grid", "QK", "QR"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert all(np.array_equal(out[0][1][key], out[1][1][key]) for key in ("QK", "QR", "rvec"))


def test_run_with_direct_launches_matches_graph_replays():
    g = grid(64)
    out = []
    for graphs in (True, False):
        c = dm.Context(g, params(), rows_hint=64)
        dm.set_loop_graphs(c, graphs)
        c.init_quench()
        n, corr = c.run(float("inf"), 150, corr_cap=150)
        out.append((c.get_state(), corr))
        c.close()
    assert np.array_equal(out[0][1], out[1][1])
    for k in ("t1grid", "QK", "QR", "dQK", "dQR", "rvec"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k


@pytest.mark.parametrize("kind", ["log_response", "rational"])
def test_run_matches_host_loop_other_forms(kind):
    """The device loop with the log-response interpolation and on a rational-stencil grid
    (wider stencil, another RHS instantiation)."""
    if kind == "rational":
        g = dm.generate_grid(64, 1e2, 5, dm.INTERP_RATIONAL, 9, 13)
        log_resp = 0
    else:
        g = grid(64)
        log_resp = 1
    print(f"device loop, {kind}: m={g.m}", flush=True)
    p = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8, log_response=log_resp)
    dev = dm.Context(g, p, rows_hint=64)
    host = dm.Context(g, p, rows_hint=64)
    for c in (dev, host):
        c.init_quench()
    n, corr = dev.run(float("inf"), 150, corr_cap=150)
    ref = host_run(host, n)
    assert n == 150 and np.array_equal(corr, ref)
    same_state(dev, host)
    dev.close()
    host.close()
