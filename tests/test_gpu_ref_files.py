"""A data.bin written by the reference's own saveSimulationStateBinary
(io_save_binary.cpp:22-116; tests/golden/ref_state64, written by
tests/golden/make_ref_files.py through oracle/_ref after 20 main-loop iterations of the
reference from the quench) loaded by dmfe_load_state and saved back by dmfe_save_state.

The state is on the L=64 poly-9 grid (spline 5, Tmax 1e2), whose stencils are well
conditioned (sum |w| <= 601 per target). On the L=32 grid of the same family single
stencils reach sum |w| ~ 1.5e8: eps sum |w| ~ 2e-8 then separates any two summation
orders of the same interpolation, and the step-size decisions that compare error norms at
rounding level stop being reproducible across implementations."""
import os
import struct

import numpy as np
import pytest

import dynamite_amd as dm
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DATA = os.path.join(GOLD, "ref_state64", "data.bin")
N = 64


def setup():
    g = dm.generate_grid(N, 1e2, 5, dm.INTERP_POLY, 9, 0)  # the grid the reference ran on
    prm = dm.default_params(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    return g, prm, np.load(os.path.join(GOLD, "ref_files.npz"))


def test_load_reference_data_bin_bit_exactly():
    g, prm, f = setup()
    ctx = dm.Context(g, prm)
    E = ctx.load_state(DATA)
    st = ctx.get_state()
    for k in ("t1grid", "QK", "QR", "dQK", "dQR", "rvec", "drvec"):
        assert np.array_equal(st[k], f[f"state_{k}"]), k
    # delta_t_ratio rebuilt from t1grid as loadSimulationStateBinary does (io_input.cpp:169-172)
    assert np.array_equal(st["dtratio"][2:], f["state_dtratio"][2:])
    c = ctx.get_control()
    dt, delta, _, loop, _ = f["state_control"]
    assert c["loop"] == int(loop) and c["delta_t"] == dt and c["delta"] == delta and c["delta_old"] == 0.0
    assert E == float(f["state_energy"])  # the energy line the reference stored
    ctx.close()


def test_save_writes_the_reference_bytes(tmp_path):
    g, prm, f = setup()
    ctx = dm.Context(g, prm)
    ctx.load_state(DATA)
    out = str(tmp_path / "data.bin")
    ctx.save_state(out)
    ctx.close()
    a, b = open(DATA, "rb").read(), open(out, "rb").read()
    assert len(a) == len(b)
    # header: int version, size_t T1, N, int p, p2, double lambda, T0, Gamma, delta, delta_t,
    # int loop, double energy; the energy is recomputed on the device (last bits may differ)
    off_e = 4 + 8 + 8 + 4 + 4 + 5 * 8 + 4
    assert a[:off_e] == b[:off_e]
    assert a[off_e + 8:] == b[off_e + 8:]
    ea, eb = struct.unpack("d", a[off_e:off_e + 8])[0], struct.unpack("d", b[off_e:off_e + 8])[0]
    # E = -t sum_j w_j f'(qK_j) qR_j (simulation_runner.cpp:334-337), summed on the device
    # in another order than the reference's loop: the two agree to 4 N eps t sum |terms|
    qk = f["state_QK"].reshape(-1, N)[-1]
    qr = f["state_QR"].reshape(-1, N)[-1]
    terms = g.integ * (1.5 * qk ** 2 + 2.0 * qk ** 3) * qr  # f'(q) for p = 3, q = 4, lambda = 0.5
    bound = 4 * N * 2.2e-16 * f["state_t1grid"][-1] * np.abs(terms).sum()
    assert abs(ea - eb) <= bound, (ea, eb, bound)


def test_resume_from_reference_state_matches_oracle():
    g, prm, f = setup()
    ctx = dm.Context(g, prm)
    ctx.load_state(DATA)
    c = ctx.get_control()
    orc = Oracle(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8)
    orc.set_history(f["state_t1grid"], f["state_dtratio"], f["state_rvec"], f["state_drvec"], f["state_QK"],
                    f["state_QR"], f["state_dQK"], f["state_dQR"])
    orc.set_control(c["delta_t"], c["delta"], 0.0, c["loop"], c["method"])
    orc.interpolate(0)
    for _ in range(30):
        assert ctx.loop_iteration() == orc.loop_iteration()
    sg, so = ctx.get_state(), orc.state()
    assert np.array_equal(sg["t1grid"], so["t1grid"])
    for k in ("QK", "QR"):
        assert np.abs(sg[k] - so[k]).max() <= 1e-10 * np.abs(so[k]).max(), k
    ctx.close()
