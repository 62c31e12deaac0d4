"""RHS kernel shapes and the autotuner that picks one.

A shape is (G lanes per B2 target entry, split column chunks per row); every shape runs
its own kernel instantiation (taps per lane per gather pass follow G and the stencil
width), so each is checked against the oracle: interpolated fields and RHS, the
log-response form, a 14-tap stencil (four taps per lane at G = 4) and a free-running
quench.  The autotuner times all shapes on a live history, keeps the fastest and
installs it for the grid class; contexts and batches created afterwards take it, a
batch stays bit-identical to contexts of the same shape, and tangents leave the primal
of a non-default shape unchanged.
"""
import numpy as np
import pytest

import dynamite_amd as dm
from oracle_lib import Oracle, synthetic_history

pytestmark = pytest.mark.gpu

_GRIDS = {}
SHAPES = [(4, 1), (4, 2), (8, 1), (8, 2), (16, 1), (16, 2), (64, 1), (64, 2)]


def grid(N, Tmax=1e3, order=9):
    key = (N, Tmax, order)
    if key not in _GRIDS:
        _GRIDS[key] = dm.generate_grid(N, Tmax, 5, dm.INTERP_POLY, order, 0)
    return _GRIDS[key]


def params(log_resp=0, Gamma=0.8, lam=0.5):
    return dm.default_params(p=3, p2=4, lambda_=lam, T0=1e50, Gamma=Gamma, log_response=log_resp)


@pytest.fixture
def restore_table():
    """The shape table is process-wide: put back the entries a test changes."""
    saved = {}

    def remember(N, m, log_resp=0):
        saved.setdefault((N, m, log_resp), dm.tuning_get(N, m, log_resp))

    yield remember
    for (N, m, lr), (G, split) in saved.items():
        dm.tuning_set(N, m, lr, G, split)


def fields_vs_oracle(g, shape, log_resp):
    st = synthetic_history(g.theta)
    ctx = dm.Context(g, params(log_resp))
    dm.set_rhs_shape(ctx, *shape)
    assert dm.rhs_shape(ctx) == shape
    orc = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8, log_resp=log_resp)
    ctx.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
    orc.set_history(st["t1grid"], st["dtratio"], st["rvec"], st["drvec"], st["QK"], st["QR"], st["dQK"], st["dQR"])
    ctx.interpolate(dm.POS_BSEARCH)
    orc.interpolate(0)
    fg, fo = ctx.fields(), orc.fields()
    errs = {k: np.abs(fg[k] - fo[k]).max() / np.abs(fo[k]).max() for k in dm.FIELD_NAMES + ("rInt",)}
    # log mode divides stencil sums by stencil sums (index_mat.cpp:52-63): 1e-9 as in
    # the main parity tests
    tol = 1e-12 if not log_resp else 1e-9
    assert max(errs.values()) <= tol, {k: f"{v:.2e}" for k, v in errs.items()}
    hK, hR = ctx.rhs()
    oK, oR = orc.rhs()
    assert np.abs(hK - oK).max() <= 1e-12 * np.abs(oK).max()
    assert np.abs(hR - oR).max() <= 1e-12 * np.abs(oR).max()
    ctx.close()


@pytest.mark.parametrize("shape", SHAPES)
def test_every_shape_matches_oracle(shape):
    print(f"shape G={shape[0]} split={shape[1]}, N=64 poly-9", flush=True)
    fields_vs_oracle(grid(64), shape, 0)


@pytest.mark.parametrize("shape", [(4, 2), (8, 1), (16, 2)])
def test_log_response_shapes_match_oracle(shape):
    print(f"log-response shape {shape}", flush=True)
    fields_vs_oracle(grid(64), shape, 1)


@pytest.mark.parametrize("shape", [(4, 1), (8, 2)])
def test_wide_stencil_shapes_match_oracle(shape):
    g = grid(64, order=13)
    assert g.m == 14  # G = 4: four taps per lane per pass
    print(f"m=14 shape {shape}", flush=True)
    fields_vs_oracle(g, shape, 0)


@pytest.mark.parametrize("shape", [(4, 2), (16, 1)])
def test_quench_with_shape_matches_oracle(shape):
    g = grid(64, 1e2)
    ctx = dm.Context(g, params())
    dm.set_rhs_shape(ctx, *shape)
    orc = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8)
    ctx.init_quench()
    orc.init_quench()
    for _ in range(120):
        assert ctx.loop_iteration() == orc.loop_iteration()
    sg, so = ctx.get_state(), orc.state()
    assert np.array_equal(sg["t1grid"], so["t1grid"])
    for k in ("QK", "QR"):
        assert np.abs(sg[k] - so[k]).max() <= 1e-10 * np.abs(so[k]).max(), k
    ctx.close()


def test_autotune_picks_and_installs_the_fastest(restore_table):
    g = grid(128, 1e4)
    restore_table(128, g.m)
    ctx = dm.Context(g, params(), rows_hint=512)
    ctx.init_quench()
    ctx.run(float("inf"), 200)
    best, table = dm.autotune(ctx, reps=5)
    print("autotune N=128:", {k: f"{v * 1e3:.1f} us" for k, v in table.items()}, "best", best, flush=True)
    assert set(table) == set(SHAPES)
    assert all(v > 0 for v in table.values())
    # the fastest, or the default (G 4, split 1 at N=128) within 2% of it
    assert table[best] == min(table.values()) or (best == (4, 1) and table[best] <= 1.02 * min(table.values()))
    assert dm.rhs_shape(ctx) == best
    assert dm.tuning_get(128, g.m, 0) == best
    ctx2 = dm.Context(g, params())
    assert dm.rhs_shape(ctx2) == best
    ctx2.close()
    # the tuned context keeps stepping like an untuned one up to rounding
    ref = dm.Context(g, params(), rows_hint=512)
    dm.set_rhs_shape(ref, 8, 1)
    ref.init_quench()
    ref.run(float("inf"), 200)
    for _ in range(20):
        ctx.loop_iteration()
        ref.loop_iteration()
    s1, s2 = ctx.get_state(), ref.get_state()
    assert np.array_equal(s1["t1grid"], s2["t1grid"])
    assert np.abs(s1["QK"] - s2["QK"]).max() <= 1e-12
    ctx.close()
    ref.close()


def test_autotune_needs_a_live_history():
    ctx = dm.Context(grid(64), params())
    ctx.init_quench()
    with pytest.raises(dm.DmfeError, match="t > 0"):
        dm.autotune(ctx, reps=2)
    ctx.close()


def test_batch_takes_tuned_shapes_and_matches_contexts(restore_table):
    gs = [grid(64, 1e2), grid(96, 1e2)]
    for g in gs:
        restore_table(g.N, g.m)
        dm.tuning_set(g.N, g.m, 0, 4, 2)
    specs = [(gs[0], 0.5, 0.8), (gs[1], 0.3, 0.5), (gs[0], 0.9, 0.2)]
    prm = [dm.default_params(p=3, p2=4, lambda_=lam, T0=1e50, Gamma=G) for _, lam, G in specs]
    b = dm.Batch([g for g, _, _ in specs], prm, rows_cap=64)
    b.init_quench()
    b.set_limits(float("inf"), 90)
    b.run(900, use_graph=True, chunk=9)  # until every instance stops at loop 90 (arena growth idles a step)
    for i, (g, _, _) in enumerate(specs):
        ctx = dm.Context(g, prm[i], rows_hint=64)
        assert dm.rhs_shape(ctx) == (4, 2)
        ctx.init_quench()
        ctx.run(float("inf"), 90)
        sc, sb = ctx.get_state(), b.get_state(i)
        for k in ("t1grid", "QK", "QR", "dQK", "dQR", "rvec", "drvec"):
            assert np.array_equal(sb[k], sc[k]), (i, k)
        ctx.close()
    b.close()


def test_tangents_leave_primal_of_shape_unchanged():
    g = grid(64, 1e2)
    out = []
    for tangents in (False, True):
        ctx = dm.Context(g, params(Gamma=0.7))
        dm.set_rhs_shape(ctx, 4, 2)
        ctx.init_quench()
        if tangents:
            ctx.tangent_init([1.0, 0.0], [0.0, 1.0])
            with pytest.raises(dm.DmfeError, match="before dmfe_tangent_init"):
                dm.set_rhs_shape(ctx, 8, 1)
        for _ in range(80):
            ctx.loop_iteration()
        out.append(ctx.get_state())
        ctx.close()
    for k in ("t1grid", "QK", "QR", "rvec"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_batch_autotune_keeps_the_state():
    """Autotuning between steps only rewrites RHS scratch: a batch tuned after 40 steps
    and one switched to the same width by hand stay bit-identical."""
    g = grid(64, 1e2)
    prm = [dm.default_params(p=3, p2=4, lambda_=lam, T0=1e50, Gamma=G) for lam, G in ((0.5, 0.8), (0.2, 0.4))]
    states = []
    for tune in (True, False):
        b = dm.Batch([g, g], prm, rows_cap=128)  # no arena growth within 80 steps
        b.init_quench()
        b.set_limits(float("inf"), 80)
        b.run(40, use_graph=False)
        if tune:
            best, table = b.autotune(reps=3)
            print("batch autotune:", table, "best", best, flush=True)
            assert set(table) == {4, 8, 16, 64}
            assert table[best] == min(table.values()) or (best == 4 and table[4] <= 1.02 * min(table.values()))
            assert b.lanes == best
        else:
            b.lanes = best
        b.run(40, use_graph=True, chunk=8)
        states.append([b.get_state(i) for i in range(2)])
        b.close()
    for i in range(2):
        for k in ("t1grid", "QK", "QR", "dQK", "dQR", "rvec"):
            assert np.array_equal(states[0][i][k], states[1][i][k]), (i, k)
