"""data.bin version-1 checkpoints and the command-line driver (dynamite_amd/dmfe_run).

The file layout follows io_save_binary.cpp:66-100 (int version, size_t T1, size_t N, int
p, int p2, double lambda, T0, Gamma, delta, delta_t, int loop, double energy, then the
history arrays, raw and unpadded); it is parsed here independently with numpy.  Resuming
follows loadSimulationStateBinary and init() (delta_old = 0, integrator from delta_t).
"""
import os
import subprocess

import numpy as np
import pytest

import dynamite_amd as dm

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "..", "dynamite_amd", "dmfe_run")
N = 64


def grid():
    return dm.generate_grid(N, 1e2, 5, dm.INTERP_POLY, 9, 0)


def params(**kw):
    base = dict(p=3, p2=4, lambda_=0.5, T0=1e50, Gamma=0.8)
    base.update(kw)
    return dm.default_params(**base)


def parse_data_bin(path):
    raw = open(path, "rb").read()
    hdr = np.dtype([("version", "<i4"), ("T1", "<u8"), ("N", "<u8"), ("p", "<i4"), ("p2", "<i4"), ("lam", "<f8"),
                    ("T0", "<f8"), ("Gamma", "<f8"), ("delta", "<f8"), ("dt", "<f8"), ("loop", "<i4"),
                    ("energy", "<f8")])
    assert hdr.itemsize == 80
    h = np.frombuffer(raw[:80], dtype=hdr)[0]
    T1, n = int(h["T1"]), int(h["N"])
    arr = np.frombuffer(raw[80:], dtype="<f8")
    assert arr.size == T1 * (3 + 4 * n)
    out, o = {}, 0
    for k, size in (("t1grid", T1), ("QK", T1 * n), ("QR", T1 * n), ("dQK", T1 * n), ("dQR", T1 * n), ("rvec", T1),
                    ("drvec", T1)):
        out[k] = arr[o:o + size]
        o += size
    return h, out


def test_save_layout_and_resume(tmp_path):
    g = grid()
    ctx = dm.Context(g, params())
    ctx.init_quench()
    for _ in range(40):
        t = ctx.loop_iteration()
    path = str(tmp_path / "data.bin")
    ctx.save_state(path)
    h, arrays = parse_data_bin(path)
    st, c = ctx.get_state(), ctx.get_control()
    assert h["version"] == 1 and h["N"] == N and h["T1"] == len(st["t1grid"])
    assert (h["p"], h["p2"], h["lam"], h["T0"], h["Gamma"]) == (3, 4, 0.5, 1e50, 0.8)
    assert h["delta"] == c["delta"] and h["dt"] == c["delta_t"] and h["loop"] == c["loop"]
    assert h["energy"] == ctx.observables(st["t1grid"][-1])[0]
    for k in arrays:
        assert np.array_equal(arrays[k], st[k]), k
    # resume into a fresh context and continue beside the original
    ctx2 = dm.Context(g, params())
    e = ctx2.load_state(path)
    assert e == h["energy"]
    c2 = ctx2.get_control()
    assert c2["delta_old"] == 0.0 and c2["loop"] == c["loop"] and c2["delta_t"] == c["delta_t"]
    st2 = ctx2.get_state()
    for k in ("t1grid", "dtratio", "QK", "QR", "dQK", "dQR", "rvec", "drvec"):
        assert np.array_equal(st[k], st2[k]), k
    ctx.set_control(c["delta_t"], c["delta"], 0.0, c["loop"], c["method"])  # resume semantics
    for _ in range(20):
        assert ctx.loop_iteration() == ctx2.loop_iteration()
    a, b = ctx.get_state(), ctx2.get_state()
    for k in ("t1grid", "QK", "QR", "rvec"):
        assert np.array_equal(a[k], b[k]), k
    ctx.close()
    ctx2.close()


def test_load_rejects_mismatches(tmp_path):
    g = grid()
    ctx = dm.Context(g, params())
    ctx.init_quench()
    ctx.loop_iteration()
    path = str(tmp_path / "data.bin")
    ctx.save_state(path)
    other = dm.Context(g, params(lambda_=0.6))
    with pytest.raises(dm.DmfeError, match="parameters"):
        other.load_state(path)
    g2 = dm.generate_grid(48, 1e2, 5, dm.INTERP_POLY, 9, 0)
    small = dm.Context(g2, params())
    with pytest.raises(dm.DmfeError, match="grid length"):
        small.load_state(path)
    trunc = str(tmp_path / "trunc.bin")
    open(trunc, "wb").write(open(path, "rb").read()[:-8])
    with pytest.raises(dm.DmfeError, match="truncated"):
        ctx.load_state(trunc)
    for c in (ctx, other, small):
        c.close()


def test_driver_run_and_resume(tmp_path):
    g = grid()
    gdir = str(tmp_path / "Grid_data" / str(N))
    g.write(gdir)
    out = str(tmp_path / "Results")
    args = [DRIVER, "-p", "3", "-q", "4", "-l", "0.5", "-G", "0.8", "-T", "inf", "-L", str(N), "-t", "1e9",
            "--grid", gdir, "-o", out]
    print("driver run", flush=True)
    r = subprocess.run(args + ["-m", "30"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    pdir = os.path.join(out, "p=3_q=4_lambda=0.50_T0=inf_G=0.80_L=64")
    corr = np.loadtxt(os.path.join(pdir, "correlation.txt"))
    assert corr.shape == (30, 3)
    # the same run through the library
    ctx = dm.Context(dm.read_grid(gdir), params())
    ctx.init_quench()
    n, ref = ctx.run(1e9, 30, corr_cap=30)
    assert n == 30
    assert np.allclose(corr, ref, rtol=0, atol=1e-13)
    txt = open(os.path.join(pdir, "params.txt")).read()
    for key in ("p = 3", "p2 = 4", "lambda = 0.5", "len = 64", "current_loop = 30", "current_method = RK54"):
        assert key in txt, key
    # every key the reference's resume check parses (io_input.cpp:344-450), present in the
    # reference's own params.txt (tests/golden/ref_state64, io_save_params.cpp) and here
    keys = lambda t: {l.split("=", 1)[0].strip() for l in t.splitlines() if "=" in l and not l.startswith("#")}  # noqa: E731
    resume = {"p", "p2", "lambda", "T0", "Gamma", "len", "delta_t_min", "delta_max", "use_serk2", "sparsify_sweeps",
              "log_response_interp", "tail_fit_enabled"}
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_state64", "params.txt")
    assert resume <= keys(open(gold).read())
    assert resume <= keys(txt), resume - keys(txt)
    # the grid package's provenance, mirrored as the reference does
    assert "grid_len = 64" in txt and "grid_interp_method = poly" in txt, txt
    h, _ = parse_data_bin(os.path.join(pdir, "data.bin"))
    assert h["loop"] == 30
    # a second invocation resumes from data.bin and appends
    r = subprocess.run(args + ["-m", "50"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "resumed" in r.stdout
    corr2 = np.loadtxt(os.path.join(pdir, "correlation.txt"))
    assert corr2.shape == (50, 3) and np.array_equal(corr2[:30], corr)
    h, _ = parse_data_bin(os.path.join(pdir, "data.bin"))
    assert h["loop"] == 50
    ctx.close()


def test_driver_interrupt_saves_and_exits_130(tmp_path):
    """SIGINT ends the run after the current segment: correlation rows up to there, the final
    data.bin of that state, exit status 130 (simulation_text_ui.cpp:30-65,
    simulation_runner.cpp:126,467).  The signal comes from the host; every kernel exits."""
    import signal
    import time

    g = grid()
    gdir = str(tmp_path / "Grid_data" / str(N))
    g.write(gdir)
    out = str(tmp_path / "Results")
    args = [DRIVER, "-p", "3", "-q", "4", "-l", "0.5", "-G", "0.8", "-T", "inf", "-L", str(N), "-t", "1e9",
            "-m", "100000000", "--grid", gdir, "-o", out]
    pdir = os.path.join(out, "p=3_q=4_lambda=0.50_T0=inf_G=0.80_L=64")
    cfile = os.path.join(pdir, "correlation.txt")
    print("driver run to be interrupted", flush=True)
    p = subprocess.Popen(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    try:
        t0 = time.time()
        while time.time() - t0 < 60:  # until the run has written a few segments
            if p.poll() is not None:
                break
            if os.path.exists(cfile) and os.path.getsize(cfile) > 2000 * 60:
                break
            time.sleep(0.2)
        assert p.poll() is None, "the run ended before the signal: " + p.stderr.read()
        p.send_signal(signal.SIGINT)
        stdout, stderr = p.communicate(timeout=60)
    finally:
        if p.poll() is None:
            p.kill()
    assert p.returncode == 130, (p.returncode, stderr)
    assert "interrupted" in stdout
    h, arrs = parse_data_bin(os.path.join(pdir, "data.bin"))
    corr = np.loadtxt(cfile)
    assert h["loop"] > 0 and corr.shape == (int(h["loop"]), 3)
    assert abs(corr[-1, 0] - arrs["t1grid"][-1]) <= 1e-12 * max(1.0, abs(corr[-1, 0]))  # last proposed t (14 decimals)
