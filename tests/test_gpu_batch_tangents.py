"""Sensitivities carried inside the batch launches (dmfe_batch_tangent_init /
dmfe_batch_tangent_observables) against single contexts carrying the same directions
(dmfe_tangent_init + dmfe_run + dmfe_tangent_observables).

The batch's stage and RHS launches run the same tangent device code (the tangent RHS
quadratures fused with the primal ones, the tangent finish before the primal finish) as
a context's device loop, with the scalar step control on the device for every instance,
so every instance must reproduce its context bit for bit: the same proposed times, the
same primal history and the same derivatives d(E, QK(t,0))/d(Gamma, lambda).  The
contexts' own tangents are checked against central finite differences and the reference
build in test_gpu_sensitivity.py.
"""
import numpy as np
import pytest

import dynamite_amd as dm

pytestmark = pytest.mark.gpu


def specs():
    # (Gamma, lambda) of four instances on one grid and one on a second grid size
    return [(64, 0.8, 0.5), (64, 0.6, 0.3), (64, 1.0, 0.9), (96, 0.7, 0.2), (64, 0.9, 0.0)]


_G = {}


def grid(N):
    if N not in _G:
        _G[N] = dm.generate_grid(N, 1e2, 5, dm.INTERP_POLY, 9, 0)
    return _G[N]


@pytest.mark.parametrize("K", [1, 2])
def test_batch_tangents_match_contexts(K):
    steps = 120
    dG, dL = ([1.0, 0.0], [0.0, 1.0]) if K == 2 else ([0.6], [-0.8])
    sp = specs()
    params = [dm.default_params(p=3, p2=4, lambda_=lam, T0=1e50, Gamma=G) for _, G, lam in sp]
    b = dm.Batch([grid(N) for N, _, _ in sp], params, rows_cap=steps + 16)
    b.init_quench()
    b.tangent_init(dG, dL)
    b.set_limits(float("inf"), steps)
    b.run(10 * steps, use_graph=True, chunk=16)
    bE, bQ = b.tangent_observables()
    print(f"K={K}: batch dE {bE.ravel()}, dQK0 {bQ.ravel()}", flush=True)
    for i, ((N, G, lam), p) in enumerate(zip(sp, params)):
        assert b.get_control(i)["loop"] == steps
        ctx = dm.Context(grid(N), p, rows_hint=steps + 16)
        ctx.init_quench()
        ctx.tangent_init(dG, dL)
        n, corr = ctx.run(float("inf"), steps, corr_cap=steps)
        assert n == steps
        sc, sb = ctx.get_state(), b.get_state(i)
        assert np.array_equal(sb["t1grid"], sc["t1grid"])
        assert np.array_equal(sb["QK"], sc["QK"]) and np.array_equal(sb["QR"], sc["QR"])
        cE, cQ = ctx.tangent_observables(corr[-1, 0])  # the proposed t of the last step
        assert np.array_equal(bE[i], cE), (i, bE[i], cE)
        assert np.array_equal(bQ[i], cQ), (i, bQ[i], cQ)
        ctx.close()
    b.close()


def test_batch_tangents_reject_a_started_batch():
    p = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    b = dm.Batch([grid(64)], [p], rows_cap=64)
    b.init_quench()
    b.set_limits(float("inf"), 3)
    b.run(10)
    with pytest.raises(dm.DmfeError):
        b.tangent_init([1.0], [0.0])
    b.close()
