"""Writes files with the reference's own writers (oracle/_ref/libdmfe_ref_strict.so, the
reference sources compiled by oracle/Makefile), for the tests that read them:

  tests/golden/ref_grid16/Grid_data/16/   a grid package: the reference generator's theta,
      int.dat, phi1/phi2, pos* and the poly order-5 stencil tables, written by
      write_all_grids, write_{A1,A2,B2}_interp_metadata and write_grid_generation_params
      (grid_io.cpp:176-328);
  tests/golden/ref_state64/data.bin, params.txt   saveSimulationStateBinary
      (io_save_binary.cpp:22-116) and saveParametersToFile (io_save_params.cpp:174-239)
      after 20 main-loop iterations of the reference from the quench (L=64 poly-9 grid,
      Tmax 1e2, p=3 q=4 lambda=0.5 Gamma=0.8 T0=inf);
  tests/golden/ref_files.npz   the arrays those files hold, as doubles.

    python tests/golden/make_ref_files.py      (needs oracle/_ref)
"""
import ctypes as C
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
os.environ["OMP_NUM_THREADS"] = "1"

import dynamite_amd as dm  # noqa: E402
from oracle_lib import REF_PATH, RefLib  # noqa: E402

_p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731


def grid_package(L, out_dir):
    N, TMAX, SPLINE, ORDER = 16, 1e2, 3, 5
    P = N * N
    theta, integ = np.zeros(N), np.zeros(N)
    pos = [np.zeros(P) for _ in range(3)]
    m = ORDER + 1
    inds, w = np.zeros(3 * P, np.int64), np.zeros(3 * P * m)
    L.ref_grid_theta(C.c_int64(N), C.c_double(TMAX), C.c_double(0.0), C.c_double(0.0), _p(theta))
    L.ref_grid_int(C.c_int(SPLINE), _p(integ))
    L.ref_grid_pos(C.c_double(TMAX), C.c_double(0.0), C.c_double(0.0), *[_p(x) for x in pos])
    for f in range(3):
        L.ref_grid_weights(C.c_int(f), C.c_int(0), C.c_int(ORDER), C.c_int(0),
                           C.c_void_p(inds.ctypes.data + 8 * f * P), C.c_void_p(w.ctypes.data + 8 * f * P * m))
    cwd = os.getcwd()
    tmp = tempfile.mkdtemp()
    try:
        os.chdir(tmp)
        L.ref_grid_write(b"16", *[_p(x) for x in pos], _p(inds), _p(w), C.c_int64(m), C.c_double(TMAX),
                         C.c_int(SPLINE), b"poly", C.c_int(ORDER), C.c_int(0), C.c_double(0.0), C.c_double(0.0))
    finally:
        os.chdir(cwd)
    dst = os.path.join(out_dir, "Grid_data", "16")
    shutil.rmtree(os.path.join(out_dir, "Grid_data"), ignore_errors=True)
    shutil.copytree(os.path.join(tmp, "Grid_data", "16"), dst)
    shutil.rmtree(tmp)
    phi2 = (theta[:, None] + (1 - theta[:, None]) * theta[None, :]).ravel()  # only for reference; read from file
    return dict(grid_theta=theta, grid_int=integ, grid_inds=inds, grid_weights=w, grid_posB2=pos[2],
                grid_spec=np.array([N, TMAX, SPLINE, ORDER, m]), grid_phi2_formula=phi2)


def state_file(out_dir):
    N = 64
    g = dm.generate_grid(N, 1e2, 5, dm.INTERP_POLY, 9, 0)
    ref = RefLib(g, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8)
    ref.init_quench()
    for _ in range(20):
        t = ref.loop_iteration()
    c = ref.control()
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, "data.bin")
    ref.L.ref_save_binary(path.encode())
    st = ref.state()
    out = {f"state_{k}": v for k, v in st.items()}
    out["state_control"] = np.array([c["delta_t"], c["delta"], c["delta_old"], c["loop"], c["method"]])
    out["state_energy"] = np.array(ref.energy(st["t1grid"][-1]))
    out["state_t"] = np.array(t)
    return out


def main():
    if not os.path.exists(REF_PATH):
        raise SystemExit("oracle/_ref/libdmfe_ref_strict.so is missing: run `make -C oracle ref` first")
    L = C.CDLL(REF_PATH)
    L.ref_grid_write.argtypes = [C.c_char_p] + [C.c_void_p] * 5 + [C.c_int64, C.c_double, C.c_int, C.c_char_p,
                                                                    C.c_int, C.c_int, C.c_double, C.c_double]
    L.ref_save_binary.argtypes = [C.c_char_p]
    arrays = grid_package(L, os.path.join(HERE, "ref_grid16"))
    arrays.update(state_file(os.path.join(HERE, "ref_state64")))
    np.savez(os.path.join(HERE, "ref_files.npz"), **arrays)
    print("wrote", sorted(os.listdir(os.path.join(HERE, "ref_grid16", "Grid_data", "16"))),
          sorted(os.listdir(os.path.join(HERE, "ref_state64"))), flush=True)


if __name__ == "__main__":
    main()
