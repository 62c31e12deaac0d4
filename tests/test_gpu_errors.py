"""Error paths of a live context: every rejected call raises with a message naming the
problem, leaves the context's state as it was, and the context keeps working.

The reference reports bad inputs by exceptions (std::invalid_argument for vectors of
unequal sizes, time_steps.cpp:21,55,136) and otherwise runs on; the C ABI returns a
status code with dmfe_last_error(), and the binding raises DmfeError (ValueError for
arrays whose sizes do not match T1 and N, checked before the C call).
"""
import numpy as np
import pytest

import dynamite_amd as dm
from oracle_lib import synthetic_history

pytestmark = pytest.mark.gpu


def make(N=32):
    g = dm.generate_grid(N, 1e2, 5, dm.INTERP_POLY, 9, 0)
    prm = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    return g, prm


def args(st):
    return (st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])


def test_calls_without_a_state_are_rejected():
    g, prm = make()
    ctx = dm.Context(g, prm)
    print("context without a state", flush=True)
    for call, match in ((ctx.rk_step, "no state"), (ctx.loop_iteration, "no state"),
                        (lambda: ctx.run(1.0, 10), "no state"), (lambda: ctx.observables(0.0), "no state"),
                        (lambda: ctx.sparsify(1e-10), "no state"), (ctx.interpolate, "no state")):
        with pytest.raises(dm.DmfeError, match=match):
            call()
    ctx.close()


def test_bad_state_and_control_leave_the_context_unchanged():
    g, prm = make()
    st = synthetic_history(g.theta)
    ctx = dm.Context(g, prm)
    ctx.set_state(*args(st))
    ctx.interpolate(dm.POS_BSEARCH)
    hK0, hR0 = ctx.rhs()
    before = ctx.get_state()
    print("bad set_state / set_control / interpolate / tangent_init arguments", flush=True)
    bad_t = st["t1grid"].copy()
    bad_t[5] = bad_t[4]  # not strictly increasing
    with pytest.raises(dm.DmfeError, match="strictly increasing"):
        ctx.set_state(bad_t, *args(st)[1:])
    nan_t = st["t1grid"].copy()
    nan_t[3] = np.nan
    with pytest.raises(dm.DmfeError, match="strictly increasing"):
        ctx.set_state(nan_t, *args(st)[1:])
    with pytest.raises(dm.DmfeError, match="T1 < 1"):
        ctx.set_state(*[np.zeros(0)] * 8)
    with pytest.raises(ValueError, match="do not match"):
        ctx.set_state(st["t1grid"], st["QK"][:-1], *args(st)[2:])
    with pytest.raises(ValueError, match="do not match"):
        ctx.set_state(*args(st)[:7], st["dtratio"][:-2])
    for method in (0, 99):
        with pytest.raises(dm.DmfeError, match="method"):
            ctx.set_control(0.01, 0.0, 0.0, 0, method)
    ctl = ctx.get_control()
    for dt, delta, loop in ((0.0, 0.0, 0), (-0.01, 0.0, 0), (np.nan, 0.0, 0), (np.inf, 0.0, 0), (0.01, np.nan, 0),
                            (0.01, 0.0, -1)):
        with pytest.raises(dm.DmfeError, match="delta_t must be finite"):
            ctx.set_control(dt, delta, 0.0, loop, dm.RK54)
    assert ctx.get_control() == ctl
    with pytest.raises(dm.DmfeError, match="bad mode"):
        ctx.interpolate(7)
    with pytest.raises(dm.DmfeError, match="K <= 2"):
        ctx.tangent_init([1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    after = ctx.get_state()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    # the context still computes the same RHS
    ctx.interpolate(dm.POS_BSEARCH)
    hK1, hR1 = ctx.rhs()
    assert np.array_equal(hK0, hK1) and np.array_equal(hR0, hR1)
    ctx.close()


def test_batch_rejects_bad_control_and_keeps_the_instance():
    g, prm = make()
    b = dm.Batch([g, g], [prm, prm], rows_cap=64)
    b.init_quench()
    b.run(5, use_graph=False)
    print("batch: bad set_control arguments", flush=True)
    c0, s0 = b.get_control(1), b.get_state(1)
    for dt, delta, loop, method in ((0.0, 0.0, 0, dm.RK54), (np.nan, 0.0, 0, dm.RK54), (0.01, np.inf, 0, dm.RK54),
                                    (0.01, 0.0, -3, dm.RK54)):
        with pytest.raises(dm.DmfeError, match="delta_t must be finite"):
            b.set_control(1, dt, delta, 0.0, loop, method)
    with pytest.raises(dm.DmfeError, match="bad argument"):
        b.set_control(1, 0.01, 0.0, 0.0, 0, 7)
    with pytest.raises(dm.DmfeError, match="bad argument"):
        b.set_control(2, 0.01, 0.0, 0.0, 0, dm.RK54)
    assert b.get_control(1) == c0
    s1 = b.get_state(1)
    for k in s0:
        assert np.array_equal(s0[k], s1[k]), k
    b.close()
