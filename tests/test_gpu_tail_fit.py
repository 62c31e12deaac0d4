"""The QK tail fit of the reference main loop (--tail-fit; tailFitBlendCPU tail_fit.cpp:46-265,
applied after update() at simulation_runner.cpp:135-141) on the GPU path, against the
oracle's restatement (tests/test_oracle_tail_fit.py pins the oracle against the reference
build).  The next step's first RHS uses the fields interpolated before the blend, as in
the reference (FSAL: no interpolation at stage 0), so the comparison runs over many steps
with blends applied: identical dt / integrator sequence and blend count, history within
1e-10."""
import numpy as np
import pytest

import dynamite_amd as dm
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

_G = {}


def grid(N=64, Tmax=1e2):
    key = (N, Tmax)
    if key not in _G:
        _G[key] = dm.generate_grid(N, Tmax, 5, dm.INTERP_POLY, 9, 0)
    return _G[key]


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def compare(ctx, o, what):
    sg, so = ctx.get_state(), o.state()
    assert np.array_equal(sg["t1grid"], so["t1grid"]), what
    assert np.array_equal(sg["dtratio"], so["dtratio"]), what
    for key in ("QK", "QR", "dQK", "dQR", "rvec", "drvec"):
        e = rel(sg[key], so[key])
        print(f"{what}: {key} rel err {e:.2e}", flush=True)
        assert e <= 1e-10, (what, key, e)


@pytest.mark.parametrize("N,Tmax", [(64, 1e2), (128, 1e3)])
def test_host_loop_with_tail_fit_matches_oracle(N, Tmax):
    g = grid(N, Tmax)
    p = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    ctx = dm.Context(g, p, rows_hint=512)
    o = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8)
    ctx.init_quench()
    o.init_quench()
    dm.set_tail_fit(ctx, True)
    o.set_tail_fit(True)
    for it in range(300):
        tg = ctx.loop_iteration()
        to = o.loop_iteration()
        assert tg == to, (it, tg, to)
        c, co = ctx.get_control(), o.control()
        for key in ("delta_t", "loop", "method"):
            assert c[key] == co[key], (it, key, c, co)
    ng, ag = dm.tail_fit_stats(ctx)
    no, ao = o.tail_fit_stats()
    print(f"N={N}: blends applied GPU {ng} oracle {no}; alpha {ag!r} {ao!r}", flush=True)
    assert ng == no and ng > 20
    assert abs(ag - ao) <= 1e-12 * abs(ao)
    compare(ctx, o, f"host loop N={N}")
    ctx.close()
    o.close()


def test_device_loop_with_tail_fit_matches_host_loop_and_oracle():
    """dmfe_run (device-side step control, replayed graphs) with the tail fit is bit-identical
    to the host-driven loop, across a maintenance (lowered cadence) and in two pieces, and
    matches the oracle."""
    g = grid()
    p = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    p.maint_every = 150
    p.delta_max = 1e-8
    a = dm.Context(g, p, rows_hint=512)
    b = dm.Context(g, p, rows_hint=512)
    o = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8, delta_max=1e-8)
    o.set_maint_every(150)
    # start above delta_t_min so that the dt halving at the maintenance keeps the loop running
    for x in (a, b):
        x.init_quench()
        k = x.get_control()
        x.set_control(4e-5, k["delta"], k["delta_old"], k["loop"], k["method"])
        dm.set_tail_fit(x, True)
    o.init_quench()
    k = o.control()
    o.set_control(4e-5, k["delta"], k["delta_old"], k["loop"], k["method"])
    o.set_tail_fit(True)
    for _ in range(260):
        a.loop_iteration()
        o.loop_iteration()
    n1, _ = b.run(1e300, 100)
    n2, _ = b.run(1e300, 260)
    assert n1 + n2 == 260
    sa, sb = a.get_state(), b.get_state()
    for key in sa:
        assert np.array_equal(sa[key], sb[key]), key
    assert a.get_control() == b.get_control()
    assert dm.tail_fit_stats(a) == dm.tail_fit_stats(b)
    assert dm.tail_fit_stats(a)[0] == o.tail_fit_stats()[0]
    compare(a, o, "device loop")
    for x in (a, b):
        x.close()
    o.close()


def test_tail_fit_off_changes_nothing():
    """Switched off again, the loop is that of a context that never had it."""
    g = grid()
    p = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    a = dm.Context(g, p, rows_hint=256)
    b = dm.Context(g, p, rows_hint=256)
    a.init_quench()
    b.init_quench()
    dm.set_tail_fit(a, True)
    dm.set_tail_fit(a, False)
    a.run(1e300, 120)
    b.run(1e300, 120)
    sa, sb = a.get_state(), b.get_state()
    for key in sa:
        assert np.array_equal(sa[key], sb[key]), key
    a.close()
    b.close()


def test_tail_fit_rejects_tangents():
    g = grid()
    p = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    ctx = dm.Context(g, p, rows_hint=64)
    ctx.init_quench()
    dm.set_tail_fit(ctx, True)
    with pytest.raises(dm.DmfeError):
        ctx.tangent_init([1.0], [0.0])
    ctx.close()
