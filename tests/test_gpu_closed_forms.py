"""The HIP path against the closed forms that pin the oracle (tests/test_oracle_pinning.py):
polynomial-history reproduction of every interpolant, quadrature closed forms, the
first-step known answer and the quench invariants - computed on the GPU through the C ABI."""
import numpy as np
import pytest

import dynamite_amd as dm
from test_oracle_pinning import check_fields, grid, poly_history

pytestmark = pytest.mark.gpu


def ctx_for(g, **kw):
    prm = dm.default_params(p=kw.get("p", 3), p2=kw.get("p2", 4), lambda_=kw.get("lam", 0.5), T0=1e50,
                            Gamma=kw.get("Gamma", 0.8))
    return dm.Context(g, prm)


@pytest.mark.parametrize("cubic", [False, True])
def test_gpu_interpolants_reproduce_polynomial_history(cubic):
    g = grid(48)
    st, funcs = poly_history(g.theta, 40, 3.0, cubic)
    ctx = ctx_for(g)
    ctx.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
    ctx.interpolate(dm.POS_BSEARCH)
    check_fields(ctx.fields(), ctx.positions(), g.theta, st, funcs, cubic, 1e-9)


def test_gpu_quadratures_closed_form():
    g = grid(128, 1e3, 5)
    lam, G = 0.5, 0.8
    ctx = ctx_for(g, lam=lam, Gamma=G)
    th = g.theta
    for t in (0.0, 1.0, 37.5):
        ctx.set_state([t], th, np.ones_like(th), np.zeros_like(th), np.zeros_like(th), [0.0], [0.0], [0.0])
        E, q0 = ctx.observables(t)
        assert abs(E - (-t)) <= 1e-12 * max(1, t)
        assert q0 == th[0]
        assert abs(ctx.rstep() - (G + t * (2 * lam + 3 * (1 - lam) + 1))) <= 1e-12 * max(1, t)


def test_gpu_first_step_known_answer():
    g = grid(64)
    G = 0.8
    ctx = ctx_for(g, Gamma=G)
    ctx.init_quench()
    hK, hR = ctx.rhs()
    want = G * (g.theta - 1)
    assert np.abs(hK - want).max() <= 2e-16 and np.abs(hR - want).max() <= 2e-16
    # dr of the first step: drvec[1] = (t1 - t0) * dr with dr = (f''(1) + f'(1)) sum w
    ctx.rk_step()
    s = ctx.get_state()
    dr = s["drvec"][1] / (s["t1grid"][1] - s["t1grid"][0])
    assert abs(dr - 12.5 * g.integ.sum()) <= 1e-13 * 12.5


def test_gpu_quench_invariants():
    g = grid(64)
    N = g.N
    ctx = ctx_for(g)
    ctx.init_quench()
    E = []
    for _ in range(300):
        t = ctx.loop_iteration()
        E.append(ctx.observables(t)[0])
    s = ctx.get_state()
    QK, QR = s["QK"].reshape(-1, N), s["QR"].reshape(-1, N)
    assert np.abs(QK[:, -1] - 1).max() <= 1e-12
    assert np.abs(QR[:, -1] - 1).max() <= 1e-12
    assert np.all(np.diff(E) <= 0)
    assert np.all(np.diff(QK[:, 0]) <= 0)
