"""dmfe_rhs_noalloc (the RHS with its scratch in the library's static device array)
against dmfe_rhs on the same state, through the C ABI: two calls in a row on different
contexts, the second on a smaller grid, results identical (same kernels, same shape)."""
import ctypes as C

import numpy as np
import pytest

import dynamite_amd as dm
from dynamite_amd.workloads import synthetic_history

pytestmark = pytest.mark.gpu


def rhs_both(N, split):
    g = dm.generate_grid(N, 1e3, 5, dm.INTERP_POLY, 9, 0)
    st = synthetic_history(g.theta)
    ctx = dm.Context(g, dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8))
    dm.set_rhs_shape(ctx, dm.rhs_shape(ctx)[0], split)
    ctx.set_state(st["t1grid"], st["QK"], st["QR"], st["dQK"], st["dQR"], st["rvec"], st["drvec"], st["dtratio"])
    ctx.interpolate(dm.POS_BSEARCH)
    hK, hR = ctx.rhs()
    nK, nR = np.full(N, np.nan), np.full(N, np.nan)
    rc = dm.lib().dmfe_rhs_noalloc(ctx._h, nK.ctypes.data_as(C.c_void_p), nR.ctypes.data_as(C.c_void_p))
    assert rc == 0, dm.lib().dmfe_last_error()
    ctx.close()
    return hK, hR, nK, nR


@pytest.mark.parametrize("first,second", [((256, 2), (64, 1))])
def test_rhs_noalloc_matches_rhs(first, second):
    for N, split in (first, second):
        print(f"dmfe_rhs_noalloc N={N} split={split}", flush=True)
        hK, hR, nK, nR = rhs_both(N, split)
        assert np.array_equal(nK, hK) and np.array_equal(nR, hR), (np.abs(nK - hK).max(), np.abs(nR - hR).max())
