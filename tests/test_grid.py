"""Grid-package generator (dmfe_grid_generate) and Grid_data/<L> file formats.

Pinned by: the reference's shipped Grid_data/512 package (tests/golden/grid512_rational.npz),
the reference's own monomial-reproduction test (test_interpolation_weights.cpp:38-66,
tolerance 5e-13), quadrature moments of int.dat, exact-node Kronecker deltas,
partition of unity, and - where oracle/_ref is built - bit-exactness against the
reference's generator functions on identical nodes.
"""
import ctypes as C
import os

import numpy as np
import pytest

import dynamite_amd as dm
from oracle_lib import REF_PATH, ref_available

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "grid512_rational.npz")


def test_shipped_grid512_fixture():
    """theta.dat / int.dat / inds*.dat of the reference's shipped Grid_data/512."""
    d = np.load(GOLDEN)
    g = dm.generate_grid(512, 1e5, 7, dm.INTERP_RATIONAL, 9, 13, 0.7, 0.7)
    # shipped values were written with %.17e from a higher-precision build; ours
    # agree to within one double ulp of theta's scale
    assert np.abs(g.theta - d["theta"]).max() < 1e-18
    assert np.abs(g.integ - d["integ"]).max() < 1e-17
    for f, k in enumerate(("indsA1", "indsA2", "indsB2")):
        assert np.array_equal(g.inds[f], d[k].astype(np.int64)), k
    assert g.m == 13


def test_theta_grid_properties():
    g = dm.generate_grid(128, 1e4)
    th = g.theta
    assert th[0] == 0.0 and th[-1] == 1.0
    assert np.all(np.diff(th) > 0)
    # symmetric arctan map: theta_i + theta_{N-1-i} = 1 (to rounding)
    assert np.abs(th + th[::-1] - 1).max() < 1e-15


@pytest.mark.parametrize("s", [3, 5, 7])
def test_quadrature_moments(s):
    """int.dat integrates the clamped degree-s spline space exactly, which contains every
    polynomial of degree <= s: sum_i w_i theta_i^k = 1/(k+1) for k <= s (integration.cpp:148-195)."""
    g = dm.generate_grid(256, 1e4, s)
    th, w = g.theta, g.integ
    for k in range(0, s + 1):
        assert abs(np.dot(w, th ** k) - 1.0 / (k + 1)) < 1e-12, k
    # and smooth functions to spline accuracy
    assert abs(np.dot(w, np.exp(th)) - (np.e - 1)) < 1e-8


def _monomial_check(g, n, tol):
    """sum_d W_kd theta_{s_k+d}^j = q_k^j for j <= n, to within tol * Lebesgue constant
    (sum_d |W_kd|) per entry: the rounding bound of evaluating the combination."""
    N, m = g.N, g.m
    th = g.theta
    phi2 = g.phi2.reshape(N, N)
    phi1 = np.outer(th, th)
    with np.errstate(divide="ignore", invalid="ignore"):
        qB2 = np.where(phi2 != 0, th[:, None] / phi2, 0.0)
    queries = (phi1.ravel(), phi2.ravel(), qB2.ravel())
    for f in range(3):
        s = np.clip(g.inds[f], 0, None)
        W = g.weights[f].reshape(-1, m)
        idx = s[:, None] + np.arange(m)[None, :]
        assert np.all(idx < N)
        nodes = th[idx]
        leb = np.abs(W).sum(axis=1)
        assert np.all(np.abs(W.sum(axis=1) - 1) <= tol * leb), f"partition of unity, table {f}"
        for d in range(1, n + 1):
            err = np.abs((W * nodes ** d).sum(axis=1) - queries[f] ** d)
            worst = (err / leb).max()
            assert worst <= tol, f"monomial theta^{d} table {f}: {worst:.2e} x Lebesgue"


def test_poly_weights_reproduce_monomials():
    g = dm.generate_grid(64, 1e3, 5, dm.INTERP_POLY, 9)
    # every table entry; the reference checks 10 equispaced queries at 5e-13 absolute
    _monomial_check(g, 9, 64 * 2.2e-16)


def test_rational_weights_reproduce_low_monomials():
    g = dm.generate_grid(64, 1e3, 5, dm.INTERP_RATIONAL, 9, 13)
    # Floater-Hormann of order d reproduces polynomials up to degree d on its window
    _monomial_check(g, 9, 64 * 2.2e-16)


def test_exact_nodes_are_kronecker():
    g = dm.generate_grid(64, 1e3)
    N, m = g.N, g.m
    th = g.theta
    # B2 queries theta_i / phi2_ij hit nodes on the diagonal block j = 0 (phi2_i0 = theta_i -> 1)
    W = g.weights[0].reshape(-1, m)
    # A1 query phi1_{0j} = 0 = theta_0: weight 1 on node 0
    for j in range(N):
        k = 0 * N + j
        assert g.inds[0][k] == 0 and W[k, 0] == 1.0 and np.all(W[k, 1:] == 0.0)
    # A2 query phi2_{i,N-1} = 1 = theta_{N-1}
    W2 = g.weights[1].reshape(-1, m)
    for i in range(N):
        k = i * N + N - 1
        assert g.inds[1][k] + m - 1 == N - 1 and W2[k, -1] == 1.0


def test_bspline_dense_weights():
    g = dm.generate_grid(40, 1e2, 3, dm.INTERP_BSPLINE, 5)
    assert g.m == 40 and np.all(g.inds == -1)
    W = g.weights[1].reshape(-1, 40)
    th = g.theta
    q = g.phi2
    # interpolating spline of degree 5 reproduces polynomials up to degree 5 (ill-conditioned dense solve)
    for d in range(0, 4):
        assert np.abs(W @ th ** d - q ** d).max() < 1e-6


def test_grid_files_round_trip(tmp_path):
    g = dm.generate_grid(32, 1e2, 3, dm.INTERP_POLY, 5, with_pos=True)
    d = str(tmp_path / "32")
    g.write(d)
    names = sorted(os.listdir(d))
    for f in ("theta.dat", "int.dat", "phi1.dat", "phi2.dat", "posA1y.dat", "posA2y.dat", "posB2y.dat",
              "indsA1y.dat", "indsA2y.dat", "indsB2y.dat", "weightsA1y.dat", "weightsA2y.dat", "weightsB2y.dat",
              "grid_params.txt"):
        assert f in names
    # formats of grid_io.cpp: one %.17e per line; N x N TSV
    first = open(os.path.join(d, "theta.dat")).readline().strip()
    assert first == "0.00000000000000000e+00"
    row = open(os.path.join(d, "indsA1y.dat")).readline().rstrip("\n").split("\t")
    assert len(row) == 32
    h = dm.read_grid(d)
    assert h.N == 32 and h.m == g.m
    assert np.array_equal(h.theta, g.theta) and np.array_equal(h.integ, g.integ)
    assert np.array_equal(h.phi2, g.phi2)
    assert np.array_equal(h.inds, g.inds) and np.array_equal(h.weights, g.weights)


def test_read_grid_missing_dir_raises(tmp_path):
    with pytest.raises(dm.DmfeError):
        dm.read_grid(str(tmp_path / "nope"))


@pytest.mark.skipif(not ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("N,Tmax,method,order,fh,s", [(48, 1e3, 0, 9, 0, 5), (64, 1e4, 1, 9, 13, 7),
                                                     (32, 1e2, 2, 5, 0, 3)])
def test_generator_bit_exact_vs_reference(N, Tmax, method, order, fh, s):
    """Same extended-precision nodes into the reference's generator functions."""
    ref = C.CDLL(REF_PATH)
    g = dm.generate_grid(N, Tmax, s, method, order, fh)
    thl = (C.c_longdouble * N)()
    assert dm.lib().dmfe_grid_theta_ld(C.c_int64(N), C.c_double(Tmax), C.c_double(0), C.c_double(0), thl) == 0
    ref.ref_grid_set_theta_ld(C.c_int64(N), thl)
    integ = np.zeros(N)
    ref.ref_grid_int(C.c_int(s), integ.ctypes.data_as(C.c_void_p))
    assert np.array_equal(integ, g.integ)
    P = N * N
    for which in range(3):
        inds = np.zeros(P, np.int64)
        w = np.zeros(P * g.m)
        ref.ref_grid_weights(C.c_int(which), C.c_int(method), C.c_int(order), C.c_int(g.m),
                             inds.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p))
        assert np.array_equal(inds, g.inds[which])
        assert np.array_equal(w, g.weights[which])


@pytest.mark.skipif(not ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("N,Tmax", [(48, 1e3), (128, 1e4)])
def test_pos_grids_agree_with_reference(N, Tmax):
    """posA1/posA2/posB2 (theta value -> fractional 1-based index) are the inverse of the
    theta map, computed here by root finding on the map; the reference generator
    (pos_grid.cpp) interpolates an oversampled table barycentrically.  Both invert the
    same map, so they agree to the reference's interpolation accuracy; the positions are
    monotone in the query value and lie in [1, N]."""
    ref = C.CDLL(REF_PATH)
    g = dm.generate_grid(N, Tmax, 5, dm.INTERP_POLY, 9, 0, with_pos=True)
    thl = (C.c_longdouble * N)()
    assert dm.lib().dmfe_grid_theta_ld(C.c_int64(N), C.c_double(Tmax), C.c_double(0), C.c_double(0), thl) == 0
    ref.ref_grid_set_theta_ld(C.c_int64(N), thl)
    P = N * N
    r = [np.zeros(P) for _ in range(3)]
    ref.ref_grid_pos(C.c_double(Tmax), C.c_double(0), C.c_double(0), *[a.ctypes.data_as(C.c_void_p) for a in r])
    for f in range(3):
        mine = np.asarray(g.pos[f]).ravel()
        assert mine.min() >= 1.0 and mine.max() <= N
        assert np.abs(mine - r[f]).max() < 1e-6, (f, np.abs(mine - r[f]).max())
    # monotone in the query: phi1 and the positions sorted together
    q = np.asarray(g.phi1).ravel()
    o = np.argsort(q, kind="stable")
    assert np.all(np.diff(np.asarray(g.pos[0]).ravel()[o]) >= 0)
