"""The oracle's SERK2 (the reference GPU loop's stabilized integrator, restated on the
CPU path's searches), pinned without a reference run:

- coefficients: sum_k b_k P_k(z) with the Chebyshev stage polynomials P_k equals
  R(z) = T_s(w0 + w1 z) / T_s(w0) as polynomials, so also off the fitting points
  z = 1..s+1 and at z = 0 (sum b = 1), evaluated with numpy's Chebyshev series;
- the step: the accepted last stage T_2(x)^q is consistent but only first-order on the
  linear part (its z^2 coefficient is (4q - 3) / (8q), not 1/2), so against an RK54
  step from the same state the difference and the error estimate shrink 4x when dt
  halves (a wrong sign or stage size would break consistency: 2x);
- its stages advance the time by exactly dt (up to rounding).
"""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as Ch

import dynamite_amd as dm
from oracle_lib import Oracle

_G = {}


def grid():
    if "g" not in _G:
        _G["g"] = dm.generate_grid(64, 1e2, 5, dm.INTERP_POLY, 9, 0)
    return _G["g"]


def T(i, x):
    return Ch.chebval(x, [0] * i + [1])


@pytest.mark.parametrize("q", [2, 4, 6])
def test_coefficients_reproduce_the_stability_polynomial(q):
    o = Oracle(grid())
    b = o.serk2_coeffs(q)
    s = 2 * q
    w0 = 1 + 0.95 / s**2
    w1 = T(s, w0) / Ch.chebval(w0, Ch.chebder([0] * s + [1]))
    for z in (-5.3, -1.0, 0.0, 0.4, 1.5, 2.5, float(s) + 0.7):
        x = 1 + z / (4 * q)
        P = [1.0] + [T(i, x) * T(2, x) ** (j - 1) for j in range(1, q + 1) for i in (1, 2)]
        R = T(s, w0 + w1 * z) / T(s, w0)
        # the system is ill-conditioned as q grows (b alternates, |b| ~ 1e2 at q = 6):
        # compare at the scale of the sum's terms
        assert abs(b @ np.array(P) - R) <= 1e-9 * np.abs(b * np.array(P)).sum(), (q, z)
    assert abs(b.sum() - 1.0) <= 1e-12
    o.close()


def test_step_is_consistent_first_order():
    g = grid()
    o = Oracle(g)
    o.init_quench()
    for _ in range(200):
        o.loop_iteration()
    st = o.state()
    o.close()
    for method in (3, 4):  # SERK2 with q = 2, 4
        diffs, ests = [], []
        for h in (1e-2, 5e-3, 2.5e-3):
            res = []
            for m in (method, 1):
                oo = Oracle(g)
                oo.set_history(st["t1grid"], st["dtratio"], st["rvec"], st["drvec"], st["QK"], st["QR"], st["dQK"],
                               st["dQR"])
                oo.interpolate(0)
                oo.set_control(h, 1.0, 0.0, 10, m)
                e = oo.update()
                s2 = oo.state()
                assert abs(s2["t1grid"][-1] - (st["t1grid"][-1] + h)) <= 1e-15 * s2["t1grid"][-1]
                res.append((s2["QK"][-64:], e))
                oo.close()
            diffs.append(np.abs(res[0][0] - res[1][0]).max())
            ests.append(res[0][1])
        for k in range(2):
            assert 3.5 < diffs[k] / diffs[k + 1] < 4.5, diffs
            assert 3.5 < ests[k] / ests[k + 1] < 4.5, ests
