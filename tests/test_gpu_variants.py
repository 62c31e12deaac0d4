"""The reference GPU path's interpolation variants (dmfe_set_gpu_variants: bisection at every
interpolation, per-element B1 clamp; interpolation_core.cu:23-42,65-72) against the
oracle with the same variants (tests/test_oracle_pinning.py shows the clamp changes no
value): host and device loops through the SSPRK104 stages after a maintenance, identical
dt sequence and bit-identical positions, history within 1e-10."""
import numpy as np
import pytest

import dynamite_amd as dm
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("device_loop", [False, True])
def test_gpu_variants_match_oracle(device_loop):
    g = dm.generate_grid(64, 1e2, 5, dm.INTERP_POLY, 9, 0)
    p = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    p.maint_every = 150
    p.delta_max = 1e-8
    ctx = dm.Context(g, p, rows_hint=512)
    o = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8, delta_max=1e-8)
    o.set_maint_every(150)
    ctx.init_quench()
    o.init_quench()
    for x in (ctx, o):
        k = x.get_control() if x is ctx else x.control()
        x.set_control(4e-5, k["delta"], k["delta_old"], k["loop"], k["method"])
    dm.set_gpu_variants(ctx, True)
    o.set_gpu_variants(True)
    methods = []
    # 200 steps across the maintenance at loop 150, then SSPRK104 forced from loop 200
    for lo, hi in ((0, 200), (200, 260)):
        if lo == 200:
            for x in (ctx, o):
                k = x.get_control() if x is ctx else x.control()
                x.set_control(k["delta_t"], k["delta"], k["delta_old"], k["loop"], dm.SSPRK104)
        if device_loop:
            for _ in range(lo, hi):
                o.loop_iteration()
                methods.append(o.control()["method"])
            n, _ = ctx.run(1e300, hi)
            assert n == hi - lo
        else:
            for it in range(lo, hi):
                tg, to = ctx.loop_iteration(), o.loop_iteration()
                assert tg == to, it
                c, co = ctx.get_control(), o.control()
                for key in ("delta_t", "loop", "method"):
                    assert c[key] == co[key], (it, key, c, co)
                methods.append(co["method"])
    print(f"SSPRK104 steps: {methods.count(dm.SSPRK104)} of {len(methods)}", flush=True)
    assert methods.count(dm.SSPRK104) >= 2 and dm.RK54 in methods  # both integrators' search sequences
    assert ctx.get_control()["delta_t"] == o.control()["delta_t"]
    sg, so = ctx.get_state(), o.state()
    assert np.array_equal(sg["t1grid"], so["t1grid"])
    for key in ("QK", "QR", "dQK", "dQR", "rvec", "drvec"):
        e = rel(sg[key], so[key])
        print(f"variants device_loop={device_loop}: {key} rel err {e:.2e}", flush=True)
        assert e <= 1e-10, key
    b1, b2 = ctx.positions()
    o1, o2 = o.positions()
    assert np.array_equal(b1, o1) and np.array_equal(b2, o2)
    ctx.close()
    o.close()
