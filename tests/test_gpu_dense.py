"""Dense B-spline grids (m = N): the FP64 MFMA contraction and the step path on it.

The contraction Y = W . X (dmfe_dense_apply) is checked against numpy's float64
product with the componentwise rounding bound K * eps * (|W| . |X|) scaled by 4, on
shapes that exercise partial row blocks (P not a multiple of 64), a partial K chunk,
odd K (the unaligned load path) and 1..16 right-hand sides.  The interpolated fields,
the RHS and a quench on a dense grid are then compared with the oracle (which reads
the intended full-row stencil; the reference's -1 start is a documented deviation).
"""
import numpy as np
import pytest

import dynamite_amd as dm
from oracle_lib import Oracle, synthetic_history

pytestmark = pytest.mark.gpu

GRIDS = {}


def bgrid(N, Tmax=1e2, order=5):
    key = (N, Tmax, order)
    if key not in GRIDS:
        GRIDS[key] = dm.generate_grid(N, Tmax, 3, dm.INTERP_BSPLINE, order, 0)
    return GRIDS[key]


def params(log_resp=0):
    return dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8, log_response=log_resp)


@pytest.mark.parametrize("N,B", [(48, 4), (40, 1), (37, 3), (64, 16), (130, 2)])
def test_dense_contraction_matches_float64_product(N, B):
    g = bgrid(N)
    assert g.m == N
    ctx = dm.Context(g, params(), rows_hint=8)
    rng = np.random.default_rng(20261203 + N)
    X = rng.uniform(-1.0, 1.0, size=(N, B))
    for map_ in (0, 1, 2):
        print(f"dense N={N} B={B} map={map_}", flush=True)
        W = g.weights[map_].reshape(N * N, N)
        Y, ms = dm.dense_apply(ctx, X, map_)
        ref = W @ X
        bound = 4 * N * 2.2e-16 * (np.abs(W) @ np.abs(X))
        assert np.all(np.abs(Y - ref) <= bound + 1e-300), np.max(np.abs(Y - ref) / (bound + 1e-300))
        assert ms > 0
    ctx.close()


@pytest.mark.parametrize("log_resp", [0, 1])
def test_dense_fields_and_rhs_match_oracle(log_resp):
    N = 48
    g = bgrid(N)
    st = synthetic_history(g.theta)
    ctx = dm.Context(g, params(log_resp))
    orc = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8, log_resp=log_resp)
    ctx.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
    orc.set_history(st["t1grid"], st["dtratio"], st["rvec"], st["drvec"], st["QK"], st["QR"], st["dQK"], st["dQR"])
    ctx.interpolate(dm.POS_BSEARCH)
    orc.interpolate(0)
    fg, fo = ctx.fields(), orc.fields()
    errs = {k: np.abs(fg[k] - fo[k]).max() / np.abs(fo[k]).max() for k in dm.FIELD_NAMES + ("rInt",)}
    tol = 1e-12 if not log_resp else 1e-9
    assert max(errs.values()) <= tol, {k: f"{v:.2e}" for k, v in errs.items()}
    hK, hR = ctx.rhs()
    oK, oR = orc.rhs()
    assert np.abs(hK - oK).max() <= 1e-12 * np.abs(oK).max()
    assert np.abs(hR - oR).max() <= 1e-12 * np.abs(oR).max()
    ctx.close()


def test_dense_quench_matches_oracle():
    N = 40
    g = bgrid(N)
    ctx = dm.Context(g, params())
    orc = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8)
    ctx.init_quench()
    orc.init_quench()
    for _ in range(60):
        tg = ctx.loop_iteration()
        to = orc.loop_iteration()
        assert tg == to
    sg, so = ctx.get_state(), orc.state()
    assert np.array_equal(sg["t1grid"], so["t1grid"])
    for key in ("QK", "QR"):
        assert np.abs(sg[key] - so[key]).max() <= 1e-11 * np.abs(so[key]).max(), key
    ctx.close()


def test_dense_degree9_fields_and_rhs_N128_match_oracle():
    """The configuration's dense family (B-spline degree 9, spline 5, Tmax 1e5) at N = 128."""
    g = dm.generate_grid(128, 1e5, 5, dm.INTERP_BSPLINE, 9, 0)
    assert g.m == 128
    st = synthetic_history(g.theta)
    ctx = dm.Context(g, params())
    orc = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8)
    ctx.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
    orc.set_history(st["t1grid"], st["dtratio"], st["rvec"], st["drvec"], st["QK"], st["QR"], st["dQK"], st["dQR"])
    ctx.interpolate(dm.POS_BSEARCH)
    orc.interpolate(0)
    fg, fo = ctx.fields(), orc.fields()
    errs = {k: np.abs(fg[k] - fo[k]).max() / np.abs(fo[k]).max() for k in dm.FIELD_NAMES + ("rInt",)}
    print("dense N=128 degree 9: fields max rel err", f"{max(errs.values()):.2e}", flush=True)
    assert max(errs.values()) <= 1e-12, {k: f"{v:.2e}" for k, v in errs.items()}
    hK, hR = ctx.rhs()
    oK, oR = orc.rhs()
    assert np.abs(hK - oK).max() <= 1e-12 * np.abs(oK).max()
    assert np.abs(hR - oR).max() <= 1e-12 * np.abs(oR).max()
    ctx.close()


def test_dense_contraction_full_size_N512():
    """The contraction at the configuration's shape: W [512^2][512] . X [512][4] (B = 4
    right-hand sides), W seeded uniform in [-1, 1] (the kernel does not depend on the
    values; the three stencil maps share the one table to keep host memory at 1 GB)."""
    from numpy.lib.stride_tricks import as_strided

    N, B = 512, 4
    g0 = dm.generate_grid(N, 1e5, 5, dm.INTERP_POLY, 9, 0)
    rng = np.random.default_rng(20261203)
    W = rng.uniform(-1.0, 1.0, size=N * N * N)
    g = dm.GridPackage.__new__(dm.GridPackage)
    g.N, g.m = N, N
    g.theta, g.integ, g.phi2 = g0.theta, g0.integ, g0.phi2
    g.inds = np.full((3, N * N), -1, np.int64)
    g.weights = as_strided(W, shape=(3, N * N * N), strides=(0, 8), writeable=False)
    g.phi1 = g.pos = g.spec = None
    ctx = dm.Context(g, params(), rows_hint=8)
    X = rng.uniform(-1.0, 1.0, size=(N, B))
    Wm = W.reshape(N * N, N)
    ref = Wm @ X
    bound = 4 * N * 2.2e-16 * (np.abs(Wm) @ np.abs(X))
    for map_ in (0, 2):
        Y, ms = dm.dense_apply(ctx, X, map_)
        err = np.max(np.abs(Y - ref) / (bound + 1e-300))
        print(f"dense N=512 B=4 map {map_}: {1e3 * ms:.1f} us, max |Y-ref| / bound = {err:.3f}", flush=True)
        assert np.all(np.abs(Y - ref) <= bound + 1e-300), err
    ctx.close()


@pytest.mark.parametrize("N,Tmax,order", [(64, 1e2, 5), (96, 1e2, 5), (128, 1e2, 7)])
def test_device_bspline_weights_match_host_solve(N, Tmax, order):
    """dmfe_grid_generate_device: the dense B-spline weights formed on the device (rows of
    the long-double inverse collocation matrix combined per query) against the host's
    per-query LU solve (interpolation_weights.cpp:401-456): every other table identical,
    the weights equal to rounding on the scale of the Lebesgue constant (max_q sum |w_q|),
    and both reproduce polynomials of the spline degree at the query points to that scale.
    The grids are well conditioned; on the configuration's family (Tmax 1e5, degree 9) the
    collocation matrix is numerically singular (Lebesgue constants ~1e36 at N = 128), in
    the reference's algorithm and here alike, and the weights of both solves are rounding
    noise there."""
    host = dm.generate_grid(N, Tmax, 5, dm.INTERP_BSPLINE, order, 0)
    dev = dm.generate_grid(N, Tmax, 5, dm.INTERP_BSPLINE, order, 0, device=True)
    assert np.array_equal(host.theta, dev.theta) and np.array_equal(host.integ, dev.integ)
    assert np.array_equal(host.phi2, dev.phi2) and np.array_equal(host.inds, dev.inds)
    P = N * N
    th = np.asarray(host.theta)
    qs = [np.asarray(host.phi1).reshape(P), np.asarray(host.phi2).reshape(P)]
    for f in range(3):
        wh = np.asarray(host.weights[f]).reshape(P, N)
        wd = np.asarray(dev.weights[f]).reshape(P, N)
        leb = np.abs(wh).sum(axis=1).max()
        err = np.abs(wd - wh).max()
        print(f"N={N} Tmax={Tmax:g} degree {order} table {f}: Lebesgue {leb:.3g}, max |dW| {err:.2e}", flush=True)
        assert err <= 1e-12 * leb, (f, err, leb)
        if f < 2:  # A1, A2 query points phi1, phi2: reproduction of x^k, k <= degree
            for k in range(order + 1):
                got = wd @ th ** k
                assert np.abs(got - qs[f] ** k).max() <= 1e-12 * leb, (f, k)
