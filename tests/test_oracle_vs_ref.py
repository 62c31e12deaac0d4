"""The C oracle against the reference's own CPU sources (oracle/_ref, built from the
reference tree by oracle/Makefile with -fno-fast-math -ffp-contract=off, one thread):
bit-identical histories, positions, fields and controls along quench trajectories."""
import numpy as np
import pytest

import dynamite_amd as dm
from oracle_lib import Oracle, RefLib, ref_available

pytestmark = [pytest.mark.ref, pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")]


def same_state(a, b):
    sa, sb = a.state(), b.state()
    return all(np.array_equal(sa[k], sb[k], equal_nan=True) for k in sa)


@pytest.mark.parametrize("log_resp", [0, 1])
def test_trajectory_bit_exact(log_resp):
    g = dm.generate_grid(40, 10.0, 5, dm.INTERP_POLY, 9, 0)
    ref = RefLib(g, Gamma=0.8, log_resp=log_resp)
    orc = Oracle(g, Gamma=0.8, log_resp=log_resp)
    ref.init_quench()
    orc.init_quench()
    for k in range(150):
        t1, t2 = ref.loop_iteration(), orc.loop_iteration()
        assert t1 == t2
        assert ref.energy(t1) == orc.energy(t2)
        if k % 50 == 49:
            assert same_state(ref, orc)
            assert ref.control() == orc.control()
            p1, p2 = ref.positions(), orc.positions()
            assert np.array_equal(p1[0], p2[0]) and np.array_equal(p1[1], p2[1])
    f1, f2 = ref.fields(), orc.fields()
    for k in f1:
        assert np.array_equal(f1[k], f2[k]), k
    # SSPRK104 branch, then sparsify + re-interpolation (simulation_runner.cpp:161-233)
    c = ref.control()
    for x in (ref, orc):
        x.set_control(c["delta_t"], c["delta"], c["delta_old"], c["loop"], 2)
    for _ in range(30):
        assert ref.loop_iteration() == orc.loop_iteration()
    assert same_state(ref, orc) and ref.control() == orc.control()
    for x in (ref, orc):
        x.sparsify(1e-13)
        x.interpolate(0)
    assert same_state(ref, orc)
    for _ in range(20):
        ref.loop_iteration()
        orc.loop_iteration()
    assert same_state(ref, orc) and ref.control() == orc.control()


def test_rational_grid_trajectory_bit_exact():
    g = dm.generate_grid(40, 10.0, 7, dm.INTERP_RATIONAL, 9, 13)
    ref = RefLib(g, p=2, p2=6, lam=0.9, Gamma=0.7)
    orc = Oracle(g, p=2, p2=6, lam=0.9, Gamma=0.7)
    ref.init_quench()
    orc.init_quench()
    for _ in range(100):
        ref.loop_iteration()
        orc.loop_iteration()
    assert same_state(ref, orc)


def test_searches_bit_exact():
    g = dm.generate_grid(32, 10.0)
    ref, orc = RefLib(g), Oracle(g)
    rng = np.random.default_rng(5)
    lst = np.concatenate([[0.0], np.cumsum(rng.uniform(0.01, 1.0, size=300))])
    el = rng.uniform(-1, lst[-1] + 1, size=4000)
    assert np.array_equal(ref.bsearch(lst, el), orc.bsearch(lst, el))
    rows = g.phi2
    inits = orc.bsearch(lst[:-7], rows * lst[-8])
    assert np.array_equal(ref.isearch(lst, rows, inits, 32), orc.isearch(lst, rows, inits, 32))
