"""The oracle's restatements of two reference main-loop behaviours outside the CPU-only
build's loop, pinned against the reference's own functions compiled from its sources
(oracle/_ref, strict build, one thread):

- the QK tail fit (tailFitBlendCPU, tail_fit.cpp:46-265) after every loop iteration: the
  reference keeps its state in function-local statics, so each run loads a private copy
  of the reference library; its 16-term window sums are `omp simd` reductions, whose
  order the compiler may change, so the trajectories agree to rounding (not bit for bit)
  while the blends are applied at the same iterations;
- rollbackState (simulation_control.cpp:14-124): truncation, delta_t, loop and the fresh
  interpolation, bit for bit.
"""
import ctypes as C
import os
import shutil
import tempfile

import numpy as np
import pytest

import dynamite_amd as dm
from oracle_lib import REF_PATH, Oracle, RefLib, ref_available

pytestmark = [pytest.mark.ref, pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")]


def private_ref(g, **kw):
    """A RefLib on a private copy of the reference library: fresh function-local statics."""
    d = tempfile.mkdtemp(prefix="dmfe_ref_")
    path = os.path.join(d, "libref_copy.so")
    shutil.copy(REF_PATH, path)
    return RefLib(g, path=path, **kw), d


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def test_tail_fit_oracle_matches_reference():
    g = dm.generate_grid(64, 1e2, 5, dm.INTERP_POLY, 9, 0)
    ref, tmp = private_ref(g, Gamma=0.8)
    try:
        orc = Oracle(g, Gamma=0.8)
        ref.init_quench()
        orc.init_quench()
        ref.L.ref_set_tail_fit(1)
        orc.set_tail_fit(True)
        applied_at = []
        last_qk = None
        for it in range(250):
            t1, t2 = ref.loop_iteration(), orc.loop_iteration()
            assert t1 == t2, it
            assert ref.control() == orc.control(), it
            n, _ = orc.tail_fit_stats()
            if n != len(applied_at):
                applied_at.append(it)
            # the blend changed the last row in both (or in neither): compare it each step
            qr, qo = ref.state()["QK"][-64:], orc.state()["QK"][-64:]
            assert rel(qo, qr) <= 1e-13, it
            last_qk = qo
        assert len(applied_at) > 20, applied_at
        sr, so = ref.state(), orc.state()
        assert np.array_equal(sr["t1grid"], so["t1grid"])
        for k in ("QK", "QR", "dQK", "dQR", "rvec", "drvec"):
            assert rel(so[k], sr[k]) <= 1e-12, k
        # the blends move the tail by ~2e-4 (a run without them), 1e9 times the agreement above
        plain = Oracle(g, Gamma=0.8)
        plain.init_quench()
        for _ in range(250):
            plain.loop_iteration()
        assert np.abs(plain.state()["QK"][-64:] - last_qk).max() > 1e-5
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def test_rollback_state_matches_reference():
    g = dm.generate_grid(40, 10.0, 5, dm.INTERP_POLY, 9, 0)
    ref = RefLib(g, Gamma=0.8)
    orc = Oracle(g, Gamma=0.8)
    ref.init_quench()
    orc.init_quench()
    for _ in range(60):
        ref.loop_iteration()
        orc.loop_iteration()
    orc.L.or_rollback_state.restype = C.c_int
    orc.L.or_rollback_state.argtypes = [C.c_void_p, C.c_int]
    for n in (10, 3):
        assert ref.L.ref_rollback_state(n) == 1
        assert orc.L.or_rollback_state(orc.h, n) == 1
        assert ref.T1 == orc.T1
        assert ref.control() == orc.control()
        sr, so = ref.state(), orc.state()
        for k in sr:
            assert np.array_equal(sr[k], so[k]), k
        p1, p2 = ref.positions(), orc.positions()
        assert np.array_equal(p1[0], p2[0]) and np.array_equal(p1[1], p2[1])
        f1, f2 = ref.fields(), orc.fields()
        for k in f1:
            assert np.array_equal(f1[k], f2[k]), k
    # too deep a rollback is refused by both, leaving the state
    T1 = ref.T1
    assert ref.L.ref_rollback_state(T1 - 1) == 0 and orc.L.or_rollback_state(orc.h, T1 - 1) == 0
    assert ref.T1 == orc.T1 == T1
