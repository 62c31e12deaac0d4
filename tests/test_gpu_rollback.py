"""The reference GPU main loop's rollback (simulation_runner.cpp:280-303, rollbackState
simulation_control.cpp:14-76), an opt-in of dmfe_set_rollback.

The CPU path never rolls back; the oracle restates the GPU loop's rollback
(or_set_rollback) and the runs are compared with it.  What the reference code fixes is
also checked directly: a step whose error exceeds 2 delta_max (forced here by a 1000x
step) drops the last min(10, T1 - 2) rows, which must equal the earlier history bit for
bit; delta_t becomes 0.5 min(max(t_last - t_prev, delta_t_min), rmax[method] / specRad),
the loop counter goes back by n and the integrator restarts as SSPRK104; the context
then continues exactly like a fresh context given the truncated state and control.
"""
import numpy as np
import pytest

import dynamite_amd as dm
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def setup():
    g = dm.generate_grid(64, 1e2, 5, dm.INTERP_POLY, 9, 0)
    p = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    return g, p


def test_rollback_drops_steps_and_restarts():
    g, p = setup()
    ctx = dm.Context(g, p, rows_hint=256)
    ctx.init_quench()
    for _ in range(80):
        ctx.loop_iteration()
    before, ctl = ctx.get_state(), ctx.get_control()
    T1 = len(before["t1grid"])
    dm.set_rollback(ctx, True)
    ctx.set_control(1000 * ctl["delta_t"], ctl["delta"], ctl["delta_old"], ctl["loop"], ctl["method"])
    t = ctx.loop_iteration()
    after, c2 = ctx.get_state(), ctx.get_control()
    assert c2["delta"] > 2 * p.delta_max  # the forced step did fail
    n = min(10, T1 + 1 - 2)
    assert len(after["t1grid"]) == T1 + 1 - n
    k = T1 + 1 - n
    for key in ("t1grid", "dtratio", "rvec", "drvec"):
        assert np.array_equal(after[key], before[key][:k]), key
    for key in ("QK", "QR", "dQK", "dQR"):
        assert np.array_equal(after[key], before[key][:k * 64]), key
    assert c2["loop"] == ctl["loop"] + 1 - n
    tl = before["t1grid"]
    last = tl[k - 1] - tl[k - 2]
    specRad = 4 * np.sqrt(3 * 2 * 0.5 + 4 * 3 * 0.5)  # 4 sqrt(f''(1)), p=3, q=4, lambda=0.5
    assert c2["delta_t"] == 0.5 * min(max(last, p.delta_t_min), p.rmax[ctl["method"] - 1] / specRad)
    assert c2["method"] == dm.SSPRK104
    assert t == tl[k - 1] + last
    # the positions were recomputed: same fields as a fresh context on the truncated state
    ref = dm.Context(g, p, rows_hint=256)
    ref.set_state(after["t1grid"], after["QK"], after["QR"], after["dQK"], after["dQR"], after["rvec"],
                  after["drvec"], after["dtratio"])
    ref.interpolate(dm.POS_BSEARCH)
    fa, fr = ctx.fields(), ref.fields()
    for name in dm.FIELD_NAMES:
        assert np.array_equal(fa[name], fr[name]), name
    # and the run goes on like the fresh context with the same control
    ref.set_control(c2["delta_t"], c2["delta"], c2["delta_old"], c2["loop"], c2["method"])
    for _ in range(25):
        ctx.loop_iteration()
        ref.loop_iteration()
    s1, s2 = ctx.get_state(), ref.get_state()
    for key in ("t1grid", "QK", "QR", "rvec"):
        assert np.array_equal(s1[key], s2[key]), key
    ctx.close()
    ref.close()


def test_rollback_off_by_default():
    g, p = setup()
    ctx = dm.Context(g, p, rows_hint=256)
    ctx.init_quench()
    for _ in range(40):
        ctx.loop_iteration()
    c = ctx.get_control()
    T1 = ctx.T1
    ctx.set_control(1000 * c["delta_t"], c["delta"], c["delta_old"], c["loop"], c["method"])
    ctx.loop_iteration()
    assert ctx.T1 == T1 + 1 and ctx.get_control()["loop"] == c["loop"] + 1
    ctx.close()


def test_rollback_run_matches_oracle():
    """Free-running host loop with rollbacks on, against the oracle with its rollback on:
    three forced error spikes (1000x steps) at different depths, each followed by 40
    iterations.  Every iteration's proposed t, dt, loop count, integrator and T1 are
    identical; the history is bit-identical in t1grid and rho and within 1e-10 in the
    history arrays."""
    g, p = setup()
    ctx = dm.Context(g, p, rows_hint=256)
    o = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8)
    dm.set_rollback(ctx, True)
    o.set_rollback(True)
    ctx.init_quench()
    o.init_quench()
    spikes = {60: 1000.0, 115: 1000.0, 170: 3000.0}
    rollbacks = 0
    for it in range(220):
        if it in spikes:
            c = ctx.get_control()
            co = o.control()
            assert c == co, (it, c, co)
            ctx.set_control(spikes[it] * c["delta_t"], c["delta"], c["delta_old"], c["loop"], c["method"])
            o.set_control(spikes[it] * co["delta_t"], co["delta"], co["delta_old"], co["loop"], co["method"])
        T1 = ctx.T1
        tg = ctx.loop_iteration()
        to = o.loop_iteration()
        c, co = ctx.get_control(), o.control()
        rollbacks += ctx.T1 < T1
        assert tg == to, (it, tg, to)
        for key in ("delta_t", "loop", "method"):
            assert c[key] == co[key], (it, key, c, co)
        assert ctx.T1 == o.T1, it
    assert rollbacks >= 3
    sg, so = ctx.get_state(), o.state()
    assert np.array_equal(sg["t1grid"], so["t1grid"]) and np.array_equal(sg["dtratio"], so["dtratio"])
    for key in ("QK", "QR", "dQK", "dQR", "rvec", "drvec"):
        err = np.abs(sg[key] - so[key]).max() / np.abs(so[key]).max()
        print(f"rollback run: {key} rel err {err:.2e}")
        assert err <= 1e-10, key
    ctx.close()
    o.close()


def host_records(ctx, iters):
    rows = []
    for _ in range(iters):
        t = ctx.loop_iteration()
        E, q0 = ctx.observables(t)
        rows.append((t, E, q0))
    return np.array(rows)


@pytest.mark.parametrize("tail_fit", [False, True])
def test_device_loop_rollback_matches_host_loop(tail_fit):
    """dmfe_run with rollbacks on (the host truncates between the device loop's graph
    replays) against the host-driven loop: two forced error spikes, correlation rows per
    iteration (the loop counter repeats after a rollback), history and control bit for
    bit; with the tail fit, whose smoothing restarts after a rollback."""
    g, p = setup()
    # the halved step after a rollback falls below the default delta_t_min = 1e-5 here,
    # which would end dmfe_run's loop (its while condition); dmfe_loop_iteration has none
    p.delta_t_min = 1e-8
    host = dm.Context(g, p, rows_hint=256)
    dev = dm.Context(g, p, rows_hint=64)  # small arena: grows during the run
    for c in (host, dev):
        dm.set_rollback(c, True)
        c.init_quench()
        k = c.get_control()
        c.set_control(1e-5, k["delta"], k["delta_old"], k["loop"], k["method"])  # the usual first step
        if tail_fit:
            dm.set_tail_fit(c, True)
    ref = [host_records(host, 60)]
    n, rec = dev.run(1e300, 60, corr_cap=60)
    assert n == 60
    got = [rec]
    for spike in (1000.0, 3000.0):
        for c in (host, dev):
            k = c.get_control()
            c.set_control(spike * k["delta_t"], k["delta"], k["delta_old"], k["loop"], k["method"])
        T1 = host.T1
        ref.append(host_records(host, 45))
        assert host.T1 < T1 + 45  # rows were dropped
        target = host.get_control()["loop"]
        n, rec = dev.run(1e300, target, corr_cap=64)
        assert n == 45, n
        got.append(rec)
    ref, got = np.vstack(ref), np.vstack(got)
    assert np.array_equal(ref, got)
    sa, sb = host.get_state(), dev.get_state()
    for key in sa:
        assert np.array_equal(sa[key], sb[key]), key
    assert host.get_control() == dev.get_control()
    if tail_fit:
        assert dm.tail_fit_stats(host) == dm.tail_fit_stats(dev)
    host.close()
    dev.close()
