"""Pins the C oracle (oracle/dmfe_oracle.c) by mathematics, independently of its own formulas.

  f, f', f'', f'''          exact rational evaluation (math_ops.cpp:7-37)
  bsearch / isearch         closed-form fractional position on lattices incl. exact nodes
                            and both ends (search_utils.cpp:13-131)
  LN3 / N / MatAll / R2     a history that is a polynomial in time (cubic, staggered
                            derivatives) and in theta (degree <= 9) is reproduced by every
                            interpolant (index_vec.cpp, index_mat.cpp)
  ConvA / rstep / energy    int.dat quadrature of polynomials: closed forms in t
  QKstep / QRstep / drstep2 first-step known answer at t = 0, T0 = inf
  whole step path           invariants C(t,t) = R(t,t) = 1, E(t) decreasing
The GPU path is checked against the same closed forms in tests/test_gpu_closed_forms.py.
"""
from fractions import Fraction

import numpy as np
import pytest

import dynamite_amd as dm
from oracle_lib import Oracle

GRIDS = {}


def grid(N, Tmax=1e2, s=5):
    key = (N, Tmax, s)
    if key not in GRIDS:
        GRIDS[key] = dm.generate_grid(N, Tmax, s, dm.INTERP_POLY, 9, 0)
    return GRIDS[key]


def exact_f(q, p, p2, lam, d):
    q, lam = Fraction(q), Fraction(lam)

    def term(k):
        c = 1
        for i in range(d):
            c *= (k - i)
        return c * q ** (k - d) if k - d >= 0 else Fraction(0)
    return float(lam * term(p) + (1 - lam) * term(p2))


@pytest.mark.parametrize("p,p2,lam", [(3, 4, 0.5), (2, 6, 0.9), (3, 12, 0.3), (4, 5, 0.25)])
def test_flambda_closed_form(p, p2, lam):
    o = Oracle(grid(16), p=p, p2=p2, lam=lam)
    for q in (0.0, 0.5, -0.5, 1.0, 0.3, 0.77, -0.9):
        for d in range(4):
            want = exact_f(q, p, p2, lam, d)
            got = o.flambda(q, d)
            assert abs(got - want) <= 8e-16 * max(1.0, abs(want)), (q, d, got, want)


def closed_position(lst, x):
    """1-based fractional position of x in the sorted list: k+1 + (x - t_k)/(t_{k+1} - t_k)."""
    lst = np.asarray(lst)
    if x <= lst[0]:
        return 1.0
    if x >= lst[-1]:
        return float(len(lst))
    k = np.searchsorted(lst, x, side="right") - 1
    return k + 1 + (x - lst[k]) / (lst[k + 1] - lst[k])


def test_bsearch_closed_form():
    o = Oracle(grid(16))
    rng = np.random.default_rng(7)
    for L in (1, 2, 3, 7, 64):
        lst = np.concatenate([[0.0], np.cumsum(rng.uniform(0.1, 2.0, size=L - 1))])
        xs = np.concatenate([lst, (lst[:-1] + lst[1:]) / 2 if L > 1 else [], np.linspace(0, lst[-1], 97)])
        got = o.bsearch(lst, xs)
        for x, gv in zip(xs, got):
            want = closed_position(lst, x)
            assert abs(gv - want) <= 4e-15 * want, (L, x, gv, want)
        # exact nodes map to exact integers
        assert np.array_equal(o.bsearch(lst, lst), np.arange(1, L + 1, dtype=float))


def test_isearch_agrees_with_closed_form():
    g = grid(32)
    N = g.N
    o = Oracle(g)
    rng = np.random.default_rng(11)
    lst = np.concatenate([[0.0], np.cumsum(rng.uniform(0.05, 1.0, size=40))])
    rows = g.phi2  # N rows of N elements, increasing along each row
    # seed with the positions for a shorter history (the situation after a step)
    inits = o.bsearch(lst[:-3], rows * lst[-4])
    got = o.isearch(lst, rows, inits, N)
    for x, gv in zip(rows * lst[-1], got):
        want = closed_position(lst, x)
        assert abs(gv - want) <= 1e-13 * want
    # exact node hits: elements equal to list entries
    el = lst[[0, 5, 17, 40]] / lst[-1]
    el = np.concatenate([el, np.ones(N - len(el))])[:N]
    el.sort()
    got = o.isearch(lst, el, np.ones(N), N)
    assert got[0] == 1.0 and got[-1] == len(lst)


# ---- polynomial history -------------------------------------------------------------
def poly_history(theta, T1, tmax, cubic):
    """Q(t, th) polynomial in t (degree 2 or 3) and theta (degree <= 9); rows at t_k with
    staggered derivatives dQ_k = (t_k - t_{k-1}) dQ/dt(t_{k-1}) (time_steps.cpp:140-157)."""
    t = np.concatenate([[0.0], np.cumsum(np.linspace(1.0, 2.0, T1 - 1))])
    t *= tmax / t[-1]
    c3 = 1.0 if cubic else 0.0

    def QK(tt, th):
        return 1 - 0.3 * tt * th ** 2 + 0.05 * tt ** 2 * th ** 5 - 0.01 * c3 * tt ** 3 * th ** 9

    def dQK(tt, th):
        return -0.3 * th ** 2 + 0.1 * tt * th ** 5 - 0.03 * c3 * tt ** 2 * th ** 9

    def QR(tt, th):
        return 0.5 + 0.2 * tt * th + 0.1 * tt ** 2 * th ** 3 + 0.02 * c3 * tt ** 3 * th ** 7

    def dQR(tt, th):
        return 0.2 * th + 0.2 * tt * th ** 3 + 0.06 * c3 * tt ** 2 * th ** 7

    def r(tt):
        return 0.7 + 0.4 * tt - 0.1 * tt ** 2 + 0.02 * c3 * tt ** 3

    def dr(tt):
        return 0.4 - 0.2 * tt + 0.06 * c3 * tt ** 2

    T, TH = np.meshgrid(t, theta, indexing="ij")
    st = dict(t1grid=t, QK=QK(T, TH).ravel(), QR=QR(T, TH).ravel(), rvec=r(t))
    dqk, dqr, drv = np.zeros_like(T), np.zeros_like(T), np.zeros(T1)
    h = np.diff(t)
    dqk[1:] = h[:, None] * dQK(T[:-1], TH[:-1])
    dqr[1:] = h[:, None] * dQR(T[:-1], TH[:-1])
    drv[1:] = h * dr(t[:-1])
    st.update(dQK=dqk.ravel(), dQR=dqr.ravel(), drvec=drv)
    rho = np.zeros(T1)
    rho[2:] = h[1:] / h[:-1]
    st["dtratio"] = rho
    return st, (QK, QR, r)


def expected_fields(theta, tl, funcs):
    QK, QR, r = funcs
    th = theta
    N = len(th)
    phi1 = np.outer(th, th)
    phi2 = th[:, None] + (1 - th[:, None]) * th[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        qB2 = np.where(phi2 > 0, th[:, None] / phi2, 0.0)
    tB1 = np.broadcast_to((th * tl)[:, None], (N, N))
    thB1 = np.broadcast_to(th[None, :], (N, N))
    return {
        "QKA1": QK(tl, phi1), "QRA1": QR(tl, phi1), "QKA2": QK(tl, phi2), "QRA2": QR(tl, phi2),
        "QKB1": QK(tB1, thB1), "QRB1": QR(tB1, thB1),
        "QKB2": QK(phi2 * tl, qB2), "QRB2": QR(phi2 * tl, qB2), "rInt": r(th * tl),
    }


def interior_masks(o_or_ctx_positions, T1, N):
    b1, b2 = o_or_ctx_positions
    F = np.clip(np.floor(b1), 1, np.ceil(b1.max()) - 1)
    b1_int = F < T1 - 1
    cap = np.floor(b2[-1] - 0.5)
    ix = np.maximum(np.minimum(np.floor(b2), cap), 1)
    b2_int = (ix < T1 - 1).reshape(N, N)
    return np.broadcast_to(b1_int[:, None], (N, N)), b2_int, b1_int


def check_fields(fields, positions, theta, st, funcs, cubic, tol):
    N = len(theta)
    T1 = len(st["t1grid"])
    exp = expected_fields(theta, st["t1grid"][-1], funcs)
    b1m, b2m, rm = interior_masks(positions, T1, N)
    for name, want in exp.items():
        got = fields[name].reshape(want.shape)
        if cubic and name in ("QKB1", "QRB1"):
            got, want = got[b1m], want[b1m]
        elif cubic and name in ("QKB2", "QRB2"):
            got, want = got[b2m], want[b2m]
        elif cubic and name == "rInt":
            got, want = got[rm], want[rm]
        err = np.abs(got - want).max() / max(1.0, np.abs(want).max())
        assert err < tol, f"{name}: {err:.2e}"


@pytest.mark.parametrize("cubic", [False, True])
def test_interpolants_reproduce_polynomial_history(cubic):
    g = grid(48)
    st, funcs = poly_history(g.theta, 40, 3.0, cubic)
    o = Oracle(g)
    o.set_history(st["t1grid"], st["dtratio"], st["rvec"], st["drvec"], st["QK"], st["QR"], st["dQK"], st["dQR"])
    o.interpolate(0)
    check_fields(o.fields(), o.positions(), g.theta, st, funcs, cubic, 1e-9)


def test_quadratures_closed_form():
    """energy and rstep on a single row qK = theta, qR = 1 with f = lam q^3 + (1-lam) q^4:
    E = -t int_0^1 f'(x) dx = -t and r = Gamma + t (2 lam + 3 (1 - lam) + 1)."""
    g = grid(128, 1e3, 5)
    lam, G = 0.5, 0.8
    o = Oracle(g, p=3, p2=4, lam=lam, Gamma=G)
    th = g.theta
    for t in (0.0, 1.0, 37.5):
        o.set_history([t], [0.0], [0.0], [0.0], th, np.ones_like(th), np.zeros_like(th), np.zeros_like(th))
        assert abs(o.energy(t) - (-t)) <= 1e-12 * max(1, t)
        assert abs(o.rstep() - (G + t * (2 * lam + 3 * (1 - lam) + 1))) <= 1e-12 * max(1, t)


def test_first_step_known_answer():
    """At t = 0 after the quench (T0 = inf) every interpolant is 1 and every quadrature
    carries a factor t = 0, so hK = hR = Gamma (theta - 1) (time_steps.cpp:36-46, 97-99) and
    dr = (f''(1) + f'(1)) sum_i w_i (time_steps.cpp:117-125)."""
    g = grid(64)
    G = 0.8
    o = Oracle(g, p=3, p2=4, lam=0.5, Gamma=G)
    o.init_quench()
    hK, hR = o.rhs()
    want = G * (g.theta - 1)
    assert np.abs(hK - want).max() <= 2e-16 and np.abs(hR - want).max() <= 2e-16
    qk = np.ones(g.N)
    dr = o.drstep2(qk, qk, hK, hR, 0.0)
    assert abs(dr - (9 + 3.5) * g.integ.sum()) <= 1e-14 * 12.5


def test_quench_invariants():
    """C(t,t) = R(t,t^-) = 1 on the theta = 1 column, and E(t) decreases after the quench."""
    g = grid(64)
    N = g.N
    o = Oracle(g, Gamma=0.8)
    o.init_quench()
    E = []
    for _ in range(300):
        t = o.loop_iteration()
        E.append(o.energy(t))
    s = o.state()
    QK, QR = s["QK"].reshape(-1, N), s["QR"].reshape(-1, N)
    assert np.abs(QK[:, -1] - 1).max() <= 1e-12
    assert np.abs(QR[:, -1] - 1).max() <= 1e-12
    assert np.all(np.diff(E) <= 0)
    assert np.all(np.diff(QK[:, 0]) <= 0)  # QK(t, 0) decays for T0 = inf


def test_gpu_path_b1_clamp_changes_no_value():
    """The reference GPU path clamps each B1 time row to [1, ceil(p) - 1] on its own
    (interpolation_core.cu:23-42) where the CPU path uses [1, ceil(max p) - 1]
    (interpolation_core.cpp:58-80).  They differ only at integer positions p = k < max p,
    where the CPU form takes row k with d = 0 and the GPU form row k - 1 with d = -1; the
    Hermite coefficients are then exactly (1, 0, 0, 0) and (0, 1, 0, 0), so both give
    history row k - 1 bit for bit.  Checked on a grid with uniform theta and a history at
    integer times, where every B1 position is an integer."""
    import copy

    N = 33
    g = copy.copy(dm.generate_grid(N, 10.0, 5, dm.INTERP_POLY, 9, 0))
    g.theta = np.arange(N, dtype=np.float64) / 32.0
    T = N  # t1grid = 0, 1, ..., 32: theta_i * t = i exactly
    rng = np.random.default_rng(5)
    t1 = np.arange(T, dtype=np.float64)
    QK = 1 + 0.1 * rng.standard_normal((T, N))
    QR = 1 + 0.1 * rng.standard_normal((T, N))
    dQK = 0.01 * rng.standard_normal((T, N))
    dQR = 0.01 * rng.standard_normal((T, N))
    r = 1 + rng.random(T)
    dr = 0.1 * rng.standard_normal(T)
    rho = np.ones(T)
    rho[:2] = 0.0
    out = []
    for variant in (False, True):
        o = Oracle(g)
        o.set_history(t1, rho, r, dr, QK.ravel(), QR.ravel(), dQK.ravel(), dQR.ravel())
        o.set_gpu_variants(variant)
        o.interpolate(0)
        pos = o.positions()[0]
        assert np.array_equal(pos, np.arange(1, N + 1, dtype=np.float64))  # all integer
        out.append(o.fields())
        o.close()
    for k in ("QKB1", "QRB1", "SigmaKB1", "SigmaRB1", "rInt"):
        assert np.array_equal(out[0][k], out[1][k]), k
    # and they are the history rows themselves: B1 target (j, i) at position j + 1 -> row j
    assert np.array_equal(out[0]["QKB1"].reshape(N, N), QK[:N, :])
