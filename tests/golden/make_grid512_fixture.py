"""Writes tests/golden/grid512_rational.npz from the reference's shipped grid package
Grid_data/512 (rational interpolation, order 9, FH window 13, spline order 7,
Tmax 1e5, alpha = delta = 0.7; see its grid_params.txt).  Data only: theta, int.dat
weights and the three stencil-start tables, as shipped.

    python tests/golden/make_grid512_fixture.py [/root/reference/Grid_data/512]
"""
import os
import sys

import numpy as np

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/Grid_data/512"
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "grid512_rational.npz")
theta = np.loadtxt(os.path.join(src, "theta.dat"))
integ = np.loadtxt(os.path.join(src, "int.dat"))
inds = [np.loadtxt(os.path.join(src, f"inds{n}y.dat"), dtype=np.int64).ravel() for n in ("A1", "A2", "B2")]
assert all(i.min() >= 0 and i.max() < 2**15 for i in inds)
params = open(os.path.join(src, "grid_params.txt")).read()
np.savez_compressed(out, theta=theta, integ=integ, indsA1=inds[0].astype(np.int16), indsA2=inds[1].astype(np.int16),
                    indsB2=inds[2].astype(np.int16), grid_params=np.array(params))
print(out, os.path.getsize(out), "bytes")
