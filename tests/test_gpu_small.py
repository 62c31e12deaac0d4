"""Parity at the smallest grids: N = 8 (poly order 5), 12, 17 and 33 (poly order 9,
spline order 3, Tmax 10), where one row is a fraction of a 64-column segment and the
stencil (m = 6 or 10) spans most of the row.  Fields and RHS on a seeded synthetic history
within 1e-12 of the oracle, and a 200-step quench on the host-driven and on the
device-driven loop with the identical dt sequence and the history within 1e-10."""
import numpy as np
import pytest

import dynamite_amd as dm
from dynamite_amd.workloads import synthetic_history
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

CASES = [(8, 5), (12, 9), (17, 9), (33, 9)]


def rel_err(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def params():
    return dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)


@pytest.mark.parametrize("N,order", CASES)
def test_small_grid_fields_rhs_and_quench(N, order):
    g = dm.generate_grid(N, 10.0, 3, dm.INTERP_POLY, order, 0)
    st = synthetic_history(g.theta, t_scale=10.0)
    ctx = dm.Context(g, params(), rows_hint=256)
    orc = Oracle(g, Gamma=0.8)
    ctx.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
    orc.set_history(st["t1grid"], st["dtratio"], st["rvec"], st["drvec"], st["QK"], st["QR"], st["dQK"], st["dQR"])
    ctx.interpolate(dm.POS_BSEARCH)
    orc.interpolate(0)
    b1, b2 = ctx.positions()
    o1, o2 = orc.positions()
    assert np.array_equal(b1, o1) and np.array_equal(b2, o2)
    fg, fo = ctx.fields(), orc.fields()
    errs = {k: rel_err(fg[k], fo[k]) for k in dm.FIELD_NAMES + ("rInt",)}
    assert max(errs.values()) <= 1e-12, {k: f"{v:.2e}" for k, v in errs.items()}
    (hK, hR), (oK, oR) = ctx.rhs(), orc.rhs()
    assert rel_err(hK, oK) <= 1e-12 and rel_err(hR, oR) <= 1e-12
    # quench: host loop and device loop against the oracle
    dev = dm.Context(g, params(), rows_hint=64)  # small arena: grows during the run
    for x in (ctx, dev, orc):
        x.init_quench()
    for it in range(200):
        tg, to = ctx.loop_iteration(), orc.loop_iteration()
        assert tg == to, (N, it)
    n, _ = dev.run(1e300, 200)
    assert n == 200
    so = orc.state()
    for c in (ctx, dev):
        sg = c.get_state()
        assert np.array_equal(sg["t1grid"], so["t1grid"])
        for key in ("QK", "QR", "dQK", "dQR", "rvec", "drvec"):
            assert rel_err(sg[key], so[key]) <= 1e-10, (N, key)
    print(f"N={N} m={g.m}: fields {max(errs.values()):.1e}, quench t = {so['t1grid'][-1]:.6g}", flush=True)
    for x in (ctx, dev, orc):
        x.close()
