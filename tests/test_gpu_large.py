"""Parity at the largest grids the library takes (N up to 2048): poly-9 grids (spline 5,
Tmax 1e5) with N = 1024 (the largest N whose RHS blocks keep the last two history rows in
LDS), N = 1100 (above it: last-interval B2 taps gathered from the history; not a multiple
of the 64-column segments) and N = 2048, on a seeded synthetic history of 256 rows.
Positions bit-exact, fields and RHS within 1e-12 of the oracle for the default RHS shape
and the largest split; one RK54 step teacher-forced at N = 1100 from a live state within
1e-11."""
import numpy as np
import pytest

import dynamite_amd as dm
from dynamite_amd.workloads import synthetic_history
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

_GRIDS = {}


def grid(N):
    if N not in _GRIDS:
        _GRIDS[N] = dm.generate_grid(N, 1e5, 5, dm.INTERP_POLY, 9, 0)
    return _GRIDS[N]


def params():
    return dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.6)


def rel_err(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def pair(N, T1=256):
    g = grid(N)
    st = synthetic_history(g.theta, T1=T1)
    ctx = dm.Context(g, params(), rows_hint=T1 + 8)
    orc = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.6)
    ctx.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
    orc.set_history(st["t1grid"], st["dtratio"], st["rvec"], st["drvec"], st["QK"], st["QR"], st["dQK"], st["dQR"])
    ctx.interpolate(dm.POS_BSEARCH)
    orc.interpolate(0)
    return g, ctx, orc


@pytest.mark.parametrize("N", [1024, 1100, 2048])
def test_fields_rhs_at_the_largest_grids(N):
    g, ctx, orc = pair(N)
    b1, b2 = ctx.positions()
    o1, o2 = orc.positions()
    assert np.array_equal(b1, o1) and np.array_equal(b2, o2)
    fo = orc.fields()
    oK, oR = orc.rhs()
    default = dm.rhs_shape(ctx)
    for split in sorted({default[1], 2 * max(1, N // 256)}):
        dm.set_rhs_shape(ctx, default[0], split)
        print(f"N={N} shape G={default[0]} split={split}", flush=True)
        fg = ctx.fields()
        errs = {k: rel_err(fg[k], fo[k]) for k in dm.FIELD_NAMES + ("rInt",)}
        print("  fields max rel err", f"{max(errs.values()):.2e}", flush=True)
        assert max(errs.values()) <= 1e-12, {k: f"{v:.2e}" for k, v in errs.items()}
        del fg
        hK, hR = ctx.rhs()
        eK, eR = rel_err(hK, oK), rel_err(hR, oR)
        print(f"  rhs rel err hK {eK:.2e} hR {eR:.2e}", flush=True)
        assert eK <= 1e-12 and eR <= 1e-12
    ctx.close()
    orc.close()


def test_rk54_step_N1100_teacher_forced():
    """From a live state (the quench after 300 GPU steps; steps on the synthetic history
    are not meaningful: its first rows have delta_t_ratio 0)."""
    g = grid(1100)
    ctx = dm.Context(g, params(), rows_hint=512)
    ctx.init_quench()
    ctx.run(float("inf"), 300)
    st, c = ctx.get_state(), ctx.get_control()
    orc = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.6)
    orc.set_history(st["t1grid"], st["dtratio"], st["rvec"], st["drvec"], st["QK"], st["QR"], st["dQK"], st["dQR"])
    ctx.interpolate(dm.POS_BSEARCH)
    orc.interpolate(0)
    for x in (ctx, orc):
        x.set_control(c["delta_t"], c["delta"], c["delta_old"], c["loop"], dm.RK54)
    eg = ctx.rk_step()
    eo = orc.update()
    print(f"N=1100 step: error norm gpu {eg:.6e} oracle {eo:.6e}", flush=True)
    assert abs(eg - eo) <= 1e-6 * abs(eo) + 4 * g.N * 2.2e-16
    sg, so = ctx.get_state(), orc.state()
    assert np.array_equal(sg["t1grid"], so["t1grid"]) and np.array_equal(sg["dtratio"], so["dtratio"])
    for key in ("QK", "QR", "dQK", "dQR", "rvec", "drvec"):
        e = rel_err(sg[key], so[key])
        print(f"  {key} rel err {e:.2e}", flush=True)
        assert e <= 1e-11, key
    ctx.close()
    orc.close()
