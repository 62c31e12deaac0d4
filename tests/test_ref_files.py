"""Files written by the reference's own writers (tests/golden/make_ref_files.py, through
oracle/_ref) read and written by this library's host code (no device needed):

  * the grid package tests/golden/ref_grid16/Grid_data/16 (write_all_grids and the
    interp-metadata writers, grid_io.cpp:176-328) read by dmfe_grid_read_dir: every
    array bit-identical to the doubles the reference wrote (tests/golden/ref_files.npz);
  * dmfe_grid_write_dir on those arrays writes the same bytes as the reference for
    theta.dat, int.dat and the inds / weights tables;
  * grid_params.txt: the keys and values of the reference's file.
"""
import os

import numpy as np
import pytest

import dynamite_amd as dm

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GDIR = os.path.join(GOLD, "ref_grid16", "Grid_data", "16")


def gold():
    return np.load(os.path.join(GOLD, "ref_files.npz"))


def test_reads_reference_grid_package_bit_exactly():
    f = gold()
    N, Tmax, spline, order, m = f["grid_spec"]
    g = dm.read_grid(GDIR)
    assert (g.N, g.m) == (int(N), int(m))
    assert np.array_equal(g.theta, f["grid_theta"])
    assert np.array_equal(g.integ, f["grid_int"])
    assert np.array_equal(g.inds.ravel(), f["grid_inds"])
    assert np.array_equal(g.weights.ravel(), f["grid_weights"])
    # phi2 = theta_i + (1 - theta_i) theta_j of the generator's long-double nodes, rounded
    assert np.abs(g.phi2 - f["grid_phi2_formula"]).max() <= 2.3e-16


@pytest.mark.parametrize("name", ["theta.dat", "int.dat", "indsA1y.dat", "indsA2y.dat", "indsB2y.dat",
                                  "weightsA1y.dat", "weightsA2y.dat", "weightsB2y.dat"])
def test_writes_the_reference_bytes(tmp_path, name):
    g = dm.read_grid(GDIR)
    out = str(tmp_path / "pkg")
    dm.GridPackage(g.N, g.m, g.theta, g.integ, g.phi2, g.inds, g.weights).write(out)
    with open(os.path.join(GDIR, name), "rb") as a, open(os.path.join(out, name), "rb") as b:
        assert a.read() == b.read(), name


def test_grid_params_file_matches_reference_keys():
    f = gold()
    N, Tmax, spline, order, m = f["grid_spec"]
    kv = {}
    with open(os.path.join(GDIR, "grid_params.txt")) as fh:
        for line in fh:
            if "=" in line and not line.startswith("#"):
                k, v = line.rstrip("\n").split("=", 1)
                kv[k] = v
    spec = dm.GridSpec(int(N), float(Tmax), 0.0, 0.0, int(spline), dm.INTERP_POLY, int(order), 0)
    g = dm.read_grid(GDIR)
    ours = {}
    for line in dm.GridPackage(g.N, g.m, g.theta, g.integ, g.phi2, g.inds, g.weights, spec=spec).params_text().splitlines():
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            ours[k] = v
    for k in ("len", "Tmax", "spline_order", "interp_method", "interp_order", "alpha", "delta", "subdir"):
        assert ours[k] == kv[k], (k, ours[k], kv[k])
