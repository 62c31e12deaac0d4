"""Parameter sensitivities d(E, QK(t,0))/d(Gamma, lambda) of the step path (forward mode).

The derivative is that of the discrete run with its time grid frozen, so it is checked
against central finite differences (h = 1e-6) of runs that replay the nominal run's dt
and integrator sequence step by step: on the GPU path itself, and on the CPU oracle (the
reference algorithm) as the reference gradient.  A thin torch.autograd.Function wraps
the run for the check through torch's autograd.
"""
import numpy as np
import pytest
import torch

import dynamite_amd as dm
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

N, STEPS, GAMMA, LAM, H = 64, 160, 0.7, 0.5, 1e-6
_GRID = {}


def grid():
    if "g" not in _GRID:
        _GRID["g"] = dm.generate_grid(N, 1e2, 5, dm.INTERP_POLY, 9, 0)
    return _GRID["g"]


def params(G=GAMMA, lam=LAM):
    return dm.default_params(p=3, p2=4, lambda_=lam, T0=1e50, Gamma=G)


def nominal_with_tangents(dG, dL):
    """Adaptive run from the quench carrying tangents; returns the control sequence."""
    ctx = dm.Context(grid(), params())
    ctx.init_quench()
    ctx.tangent_init(dG, dL)
    seq = []
    for _ in range(STEPS):
        c = ctx.get_control()
        seq.append((c["delta_t"], c["method"]))
        t = ctx.loop_iteration()
    E, q0 = ctx.observables(t)
    dE, dQ = ctx.tangent_observables(t)
    st = ctx.get_state()
    ctx.close()
    return seq, t, (E, q0), (dE, dQ), st


def replay_gpu(seq, t, G, lam):
    ctx = dm.Context(grid(), params(G, lam))
    ctx.init_quench()
    for dt, method in seq:
        ctx.set_control(dt, 1.0, 0.0, 10, method)
        ctx.rk_step()
    E, q0 = ctx.observables(t)
    ctx.close()
    return np.array([E, q0])


def replay_cpu(seq, t, G, lam):
    orc = Oracle(grid(), p=3, p2=4, lam=lam, T0=1e50, Gamma=G)
    orc.init_quench()
    for dt, method in seq:
        orc.set_control(dt, 1.0, 0.0, 10, method)
        orc.update()
    E = orc.energy(t)
    s = orc.state()
    q0 = s["QK"][(len(s["t1grid"]) - 1) * N]
    return np.array([E, q0])


def fd(replay, seq, t):
    """columns: d/dGamma, d/dlambda of (E, QK0)"""
    J = np.zeros((2, 2))
    for col, (dg, dl) in enumerate(((H, 0.0), (0.0, H))):
        J[:, col] = (replay(seq, t, GAMMA + dg, LAM + dl) - replay(seq, t, GAMMA - dg, LAM - dl)) / (2 * H)
    return J


@pytest.fixture(scope="module")
def nominal():
    print(f"sensitivities: N={N}, {STEPS} steps, Gamma={GAMMA}, lambda={LAM}", flush=True)
    return nominal_with_tangents([1.0, 0.0], [0.0, 1.0])


def jac(nom):
    (dE, dQ) = nom[3]
    return np.array([[dE[0], dE[1]], [dQ[0], dQ[1]]])


def test_tangent_run_leaves_primal_unchanged(nominal):
    seq, t, (E, q0), _, st = nominal
    ctx = dm.Context(grid(), params())
    ctx.init_quench()
    for _ in range(STEPS):
        ctx.loop_iteration()
    s2 = ctx.get_state()
    ctx.close()
    for k in ("t1grid", "QK", "QR", "rvec"):
        assert np.array_equal(st[k], s2[k]), k


def test_tangent_run_leaves_primal_unchanged_L256():
    """The same on the sensitivity workload's grid (L=256 poly 9, Tmax 1e5) through dmfe_run,
    far enough that B2 targets lie both inside the history and in its last interval."""
    g = dm.generate_grid(256, 1e5, 5, dm.INTERP_POLY, 9, 0)
    states = []
    for tangents in (True, False):
        ctx = dm.Context(g, params(), rows_hint=700)
        ctx.init_quench()
        if tangents:
            ctx.tangent_init([1.0, 0.0], [0.0, 1.0])
        ctx.run(float("inf"), 600)
        states.append(ctx.get_state())
        ctx.close()
    for k in ("t1grid", "QK", "QR", "dQK", "dQR", "rvec", "drvec"):
        assert np.array_equal(states[0][k], states[1][k]), k


def test_matches_gpu_finite_differences(nominal):
    seq, t = nominal[0], nominal[1]
    J, Jfd = jac(nominal), fd(replay_gpu, seq, t)
    print("tangent", J.ravel(), "fd", Jfd.ravel(), flush=True)
    assert np.all(np.abs(J - Jfd) <= 1e-6 * np.abs(Jfd).max() + 1e-9), (J, Jfd)


def test_matches_cpu_reference_gradient(nominal):
    seq, t = nominal[0], nominal[1]
    J, Jcpu = jac(nominal), fd(replay_cpu, seq, t)
    assert np.all(np.abs(J - Jcpu) <= 1e-6 * np.abs(Jcpu).max() + 1e-9), (J, Jcpu)


def test_tangent_is_linear_in_the_seed(nominal):
    J = jac(nominal)
    _, _, _, (dE, dQ), _ = nominal_with_tangents([2.0], [-3.0])
    assert abs(dE[0] - (2 * J[0, 0] - 3 * J[0, 1])) <= 1e-10 * np.abs(J[0]).sum()
    assert abs(dQ[0] - (2 * J[1, 0] - 3 * J[1, 1])) <= 1e-10 * np.abs(J[1]).sum()


class QuenchObservables(torch.autograd.Function):
    """(Gamma, lambda) -> (E, QK(t,0)) of the quench run; backward through dmfe_adjoint
    with the upstream gradient as its seeds."""

    @staticmethod
    def forward(ctx_, G, lam):
        c = dm.Context(grid(), params(float(G), float(lam)))
        c.init_quench()
        _, corr = c.run(1e300, STEPS, corr_cap=STEPS)
        E, q0 = c.observables(corr[-1][0])  # at the last proposed time
        c.close()
        ctx_.save_for_backward(G, lam)
        return torch.tensor([E, q0], dtype=torch.float64)

    @staticmethod
    def backward(ctx_, g):
        G, lam = ctx_.saved_tensors
        c = dm.Context(grid(), params(float(G), float(lam)))
        dG, dL, _, _ = c.adjoint(g.detach().cpu().numpy(), 1e300, STEPS)
        c.close()
        return torch.tensor(dG, dtype=torch.float64), torch.tensor(dL, dtype=torch.float64)


def test_adjoint_is_the_seeded_jacobian(nominal):
    """dmfe_adjoint(seeds) = seeds . J, J from dmfe_gradients on the same run."""
    c = dm.Context(grid(), params())
    grad, t, n = c.gradients(1e300, STEPS)
    c.close()
    J = grad.reshape(2, 2)
    c = dm.Context(grid(), params())
    dG, dL, t2, n2 = c.adjoint([1.0, 3.0], 1e300, STEPS)
    c.close()
    assert (t2, n2) == (t, n)
    assert dG == 1.0 * J[0, 0] + 3.0 * J[1, 0] and dL == 1.0 * J[0, 1] + 3.0 * J[1, 1]
    with pytest.raises(dm.DmfeError):
        dm.Context(grid(), params()).adjoint([np.nan, 1.0], 1e300, STEPS)


def test_torch_autograd_binding(nominal):
    seq, t = nominal[0], nominal[1]
    G = torch.tensor(GAMMA, dtype=torch.float64, requires_grad=True)
    lam = torch.tensor(LAM, dtype=torch.float64, requires_grad=True)
    out = QuenchObservables.apply(G, lam)
    (out[0] + 3.0 * out[1]).backward()
    Jfd = fd(replay_gpu, seq, t)
    want = Jfd[0] + 3.0 * Jfd[1]
    got = np.array([G.grad.item(), lam.grad.item()])
    assert np.all(np.abs(got - want) <= 1e-6 * np.abs(want).max() + 1e-9), (got, want)


def test_gradients_many_matches_single_points():
    """Several (Gamma, lambda) points at once (one context, stream and host thread each)
    give exactly the gradients of one context per point."""
    pts = [(0.7, 0.5), (0.4, 0.2), (0.9, 0.8), (0.55, 0.35), (0.3, 0.6)]
    prm = [params(G, lam) for G, lam in pts]
    grads, tf, st = dm.gradients_many([grid()] * len(pts), prm, 1e300, STEPS)
    for i, p in enumerate(prm):
        c = dm.Context(grid(), p)
        g1, t1, n1 = c.gradients(1e300, STEPS)
        c.close()
        assert np.array_equal(grads[i], g1), (i, grads[i], g1)
        assert tf[i] == t1 and st[i] == n1


def test_tangents_match_reference_build_gradient_L256():
    """The sensitivity workload's grid (L=256 poly 9, Tmax 1e5) and model (p=3+4, lambda
    0.5, Gamma 0.7) from the quench to t ~ 100: forward-mode tangents of the GPU path,
    replaying the reference's own step sequence, against the gradient the reference CPU
    path gives by central finite differences (h = 1e-6) of runs of the reference sources
    replaying that sequence (tests/golden/ref_gradient256.npz, written by
    tests/golden/make_ref_gradient256.py from oracle/_ref).  Bars: the nominal energy and
    QK(t,0) within 1e-10 relative, the Jacobian within 1e-6 of its largest entry."""
    import os

    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_gradient256.npz"))
    Ng, Tg, spline, order, p, p2, lam, G, T0, h = f["params"]
    g = dm.generate_grid(int(Ng), Tg, int(spline), dm.INTERP_POLY, int(order), 0)
    ctx = dm.Context(g, dm.default_params(p=int(p), p2=int(p2), lambda_=lam, T0=T0, Gamma=G),
                     rows_hint=len(f["dt"]) + 64)
    ctx.init_quench()
    ctx.tangent_init([1.0, 0.0], [0.0, 1.0])
    print(f"L={int(Ng)}: replaying {len(f['dt'])} reference steps with tangents", flush=True)
    for dt, method in zip(f["dt"], f["method"]):
        ctx.set_control(float(dt), 1.0, 0.0, 10, int(method))
        ctx.rk_step()
    t = float(f["t_final"])
    E, q0 = ctx.observables(t)
    dE, dQ = ctx.tangent_observables(t)
    ctx.close()
    E0, Q0 = f["observables"]
    assert abs(E - E0) <= 1e-10 * abs(E0) and abs(q0 - Q0) <= 1e-10 * abs(Q0), (E, E0, q0, Q0)
    J = np.array([[dE[0], dE[1]], [dQ[0], dQ[1]]])
    Jref = f["jacobian"]
    err = np.abs(J - Jref).max() / np.abs(Jref).max()
    print("J gpu", J.tolist(), "J reference FD", Jref.tolist(), f"max rel err {err:.2e}", flush=True)
    assert err <= 1e-6
