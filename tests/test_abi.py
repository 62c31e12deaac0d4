"""The C ABI library: loads on a machine without a GPU, exports every function that
include/*.h declares, reports errors through status codes and dmfe_last_error()."""
import ctypes as C
import glob
import os
import re

import pytest

import dynamite_amd as dm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(dmfe_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("dmfe_ctx_create", "dmfe_rk_step", "dmfe_interpolate", "dmfe_grid_generate", "dmfe_run"):
        assert must in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = dm.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_params_default_matches_reference_config():
    p = dm.default_params()
    assert (p.p, p.p2) == (3, 12) and p.lambda_ == 0.3 and p.T0 == 1e50 and p.Gamma == 0.0
    assert p.delta_t_min == 1e-5 and p.delta_max == 1e-10
    assert list(p.rmax) == [3, 13, 20, 80, 180, 320, 500, 720, 980, 1280, 1620]


def test_error_paths_without_device_work():
    lib = dm.lib()
    assert lib.dmfe_ctx_create(None, None, C.c_int64(0), None) == -1
    assert b"null" in lib.dmfe_last_error()
    assert lib.dmfe_grid_stencil_width(C.c_int64(64), 0, 9, 0) == 10
    assert lib.dmfe_grid_stencil_width(C.c_int64(64), 1, 9, 13) == 13
    assert lib.dmfe_grid_stencil_width(C.c_int64(64), 2, 9, 0) == 64
    spec = dm.GridSpec(1, 1e3, 0, 0, 5, 0, 9, 0)
    assert lib.dmfe_grid_generate(C.byref(spec), None, None, None, None, None, None, None, None, None) == -1


def test_bad_grid_rejected_before_device_use():
    g = dm.generate_grid(16, 10.0, 3, dm.INTERP_POLY, 5)
    g.inds[0][5] = 14  # window 14..19 exceeds N = 16
    with pytest.raises(dm.DmfeError, match="exceeds N"):
        dm.Context(g, dm.default_params())


def test_driver_help_runs_without_a_device():
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "dynamite_amd", "dmfe_run")
    r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--grid" in r.stdout


def test_shape_table_validates_and_round_trips():
    """The RHS shape table needs no device: defaults by stencil width, validation of
    G against m and of split against N, and set/get of a tuned entry."""
    assert dm.tuning_get(512, 10, 0) == (4, 2)
    assert dm.tuning_get(128, 10, 0) == (4, 1)
    assert dm.tuning_get(512, 20, 0) == (16, 2)
    assert dm.tuning_get(512, 512, 0) == (64, 2)
    with pytest.raises(dm.DmfeError, match="G must be"):
        dm.tuning_set(512, 20, 0, 8, 2)  # 8 lanes need m <= 16
    with pytest.raises(dm.DmfeError, match="G must be"):
        dm.tuning_set(512, 10, 0, 5, 2)
    with pytest.raises(dm.DmfeError, match="split must be"):
        dm.tuning_set(512, 10, 0, 4, 5)  # at most 2 * 512 / 256
    dm.tuning_set(300, 10, 1, 4, 2)
    assert dm.tuning_get(300, 10, 1) == (4, 2)
    assert dm.tuning_get(300, 10, 0) == (4, 1)  # another class keeps the default
    dm.tuning_set(300, 10, 1, 8, 1)
