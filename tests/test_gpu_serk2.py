"""SERK2 on the GPU path against the oracle's SERK2 (tests/test_oracle_serk2.py pins the
oracle): teacher-forced steps at two levels, and the main loop with the SERK2
promotion at a lowered maintenance cadence (RK54 -> SSPRK104 -> SERK2(2) -> SERK2(4))."""
import numpy as np
import pytest

import dynamite_amd as dm
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

_G = {}


def grid():
    if "g" not in _G:
        _G["g"] = dm.generate_grid(64, 1e2, 5, dm.INTERP_POLY, 9, 0)
    return _G["g"]


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("method", [3, 4])
def test_serk2_steps_match_oracle(method):
    g = grid()
    o = Oracle(g)
    o.init_quench()
    for _ in range(150):
        o.loop_iteration()
    st = o.state()
    ctx = dm.Context(g, dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8))
    ctx.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
    ctx.interpolate(dm.POS_BSEARCH)
    o.interpolate(0)
    dt = 2e-3
    for k in range(6):
        print(f"SERK2 method {method} step {k}", flush=True)
        ctx.set_control(dt, 1.0, 0.0, 10, method)
        o.set_control(dt, 1.0, 0.0, 10, method)
        eg = ctx.rk_step()
        eo = o.update()
        sg, so = ctx.get_state(), o.state()
        assert np.array_equal(sg["t1grid"], so["t1grid"])
        for key in ("QK", "QR", "dQK", "dQR", "rvec", "drvec"):
            assert rel(sg[key], so[key]) <= 1e-11, key
        assert abs(eg - eo) <= 1e-6 * abs(eo) + 4 * 64 * 2.2e-16, (eg, eo)
    ctx.close()
    o.close()


def test_main_loop_promotes_to_serk2_like_the_oracle():
    g = grid()
    p = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    p.maint_every = 60
    p.delta_max = 1e-3  # loose: SSPRK104 is not switched back to RK54 between maintenances
    ctx = dm.Context(g, p, rows_hint=64)
    dm.set_serk2(ctx, True)
    o = Oracle(g, delta_max=1e-3)
    o.set_serk2(True)
    o.set_maint_every(60)
    ctx.init_quench()
    o.init_quench()
    c = ctx.get_control()
    ctx.set_control(4e-5, c["delta"], c["delta_old"], c["loop"], c["method"])
    o.set_control(4e-5, c["delta"], c["delta_old"], c["loop"], c["method"])
    methods = []
    for k in range(200):
        tg = ctx.loop_iteration()
        to = o.loop_iteration()
        assert tg == to, k
        mg, mo = ctx.get_control()["method"], o.control()["method"]
        assert mg == mo, (k, mg, mo)
        methods.append(mg)
    # RK54 until loop 60, SSPRK104 until 120, SERK2(2) until 180, then SERK2(4)
    assert methods[58] == 1 and methods[59] == 2 and methods[119] == 3 and methods[179] == 4, sorted(set(methods))
    sg, so = ctx.get_state(), o.state()
    assert np.array_equal(sg["t1grid"], so["t1grid"])
    for key in ("QK", "QR"):
        assert rel(sg[key], so[key]) <= 1e-9, key
    ctx.close()
    o.close()


@pytest.mark.parametrize("graphs", [True, False])
def test_device_loop_serk2_matches_host_loop_and_oracle(graphs):
    """dmfe_run's device-driven loop through every SERK2 level: the maintenance every 60
    loops promotes RK54 -> SSPRK104 -> SERK2(2) -> ... -> SERK2(18) (method 11, 36 stages,
    the last level of the rmax table) over 660 loops.  The device loop is bit-identical to
    the host-driven loop (history, control, per-iteration t), and both follow the oracle
    with the identical t sequence."""
    g = grid()
    p = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    p.maint_every = 60
    p.delta_max = 1e-3
    p.delta_t_min = 1e-12  # each maintenance halves delta_t; dmfe_run stops below delta_t_min
    L = 660
    host, dev = dm.Context(g, p, rows_hint=64), dm.Context(g, p, rows_hint=64)
    o = Oracle(g, delta_max=1e-3, dt_min=1e-12)
    o.set_serk2(True)
    o.set_maint_every(60)
    for x in (host, dev):
        dm.set_serk2(x, True)
        x.init_quench()
    o.init_quench()
    dm.set_loop_graphs(dev, graphs)
    c = host.get_control()
    for x in (host, dev, o):
        x.set_control(4e-5, c["delta"], c["delta_old"], c["loop"], c["method"])
    th, to, methods = [], [], []
    for k in range(L):
        th.append(host.loop_iteration())
        to.append(o.loop_iteration())
        methods.append(host.get_control()["method"])
    print(f"methods reached: {sorted(set(methods))}", flush=True)
    assert methods[-1] == 11 and methods.count(11) >= 50
    assert th == to
    n, corr = dev.run(1e300, L, corr_cap=L)
    assert n == L
    assert np.array_equal(corr[:, 0], np.asarray(th))
    ch, cd = host.get_control(), dev.get_control()
    for key in ("delta_t", "delta", "delta_old", "loop", "method"):
        assert ch[key] == cd[key], (key, ch[key], cd[key])
    sh, sd, so = host.get_state(), dev.get_state(), o.state()
    for key in sh:
        assert np.array_equal(sh[key], sd[key]), key
    assert np.array_equal(sd["t1grid"], so["t1grid"])
    for key in ("QK", "QR", "dQK", "dQR", "rvec", "drvec"):
        e = rel(sd[key], so[key])
        print(f"device loop vs oracle after {L} loops: {key} rel err {e:.2e}", flush=True)
        assert e <= 1e-10, key
    for x in (host, dev, o):
        x.close()


def test_device_loop_serk2_at_L512_matches_oracle():
    """SERK2 on dmfe_run's device loop at the configuration's grid (L=512 poly-9, spline 5,
    Tmax 1e5; p=3+4, lambda=0.5, Gamma=0.6): from a live quench state (300 GPU steps,
    T1=301, taken into a context with delta_t_min 1e-12), 12 loops at SERK2(2) and 6 at
    SERK2(6) against the oracle's main loop on the same state: identical t sequence and
    control, history within 1e-10."""
    g = dm.generate_grid(512, 1e5, 5, dm.INTERP_POLY, 9, 0)
    prm = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.6)
    ctx = dm.Context(g, prm, rows_hint=400)
    ctx.init_quench()
    ctx.run(float("inf"), 300)
    # the SERK2 steps' error norms exceed 2 delta_max on this early state, so delta_t
    # shrinks by 0.9 a step; a lower delta_t_min keeps dmfe_run's loop condition open
    p2 = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.6)
    p2.delta_t_min = 1e-12
    st, c = ctx.get_state(), ctx.get_control()
    ctx.close()
    ctx = dm.Context(g, p2, rows_hint=400)
    ctx.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
    ctx.set_control(c["delta_t"], c["delta"], c["delta_old"], c["loop"], c["method"])
    o = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.6, dt_min=1e-12)
    o.set_history(st["t1grid"], st["dtratio"], st["rvec"], st["drvec"], st["QK"], st["QR"], st["dQK"], st["dQR"])
    ctx.interpolate(dm.POS_BSEARCH)
    o.interpolate(0)
    for method, loops in ((3, 12), (5, 6)):
        k = ctx.get_control()
        for x in (ctx, o):
            x.set_control(k["delta_t"], k["delta"], k["delta_old"], k["loop"], method)
        to = [o.loop_iteration() for _ in range(loops)]
        n, corr = ctx.run(float("inf"), k["loop"] + loops, corr_cap=loops)
        assert n == loops
        assert np.array_equal(corr[:, 0], np.asarray(to)), (method, corr[:, 0], to)
        cg, co = ctx.get_control(), o.control()
        for key in ("delta_t", "loop", "method"):
            assert cg[key] == co[key], (method, key, cg[key], co[key])
        sg, so = ctx.get_state(), o.state()
        assert np.array_equal(sg["t1grid"], so["t1grid"])
        for key in ("QK", "QR", "dQK", "dQR", "rvec", "drvec"):
            e = rel(sg[key], so[key])
            print(f"L=512 SERK2 method {method}: {key} rel err {e:.2e}", flush=True)
            assert e <= 1e-10, (method, key)
    ctx.close()
    o.close()
