"""Builds dynamite_amd/libdmfe.so (gfx950 HIP kernels + C++ host library).

    python -m dynamite_amd.build          # or: python dynamite_amd/build.py

Compiles in-tree with hipcc for gfx950 only.  Objects go to dynamite_amd/_build/,
the shared library to dynamite_amd/libdmfe.so, the command-line driver (g++ over the C
ABI) to dynamite_amd/dmfe_run.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdmfe.so")
DRIVER = os.path.join(HERE, "dmfe_run")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", f"-I{os.path.join(HERE, '..', 'include')}"]
# packed FP32 instructions stay disabled in device code
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
SOURCES = ["dmfe_rhs.hip", "dmfe_loop.hip", "dmfe_kernels.hip", "dmfe_ctx.cpp", "dmfe_grid.cpp", "dmfe_batch.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def build(verbose=False):
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC, *COMMON, *DEVICE, "-c", "-o", obj, path]
        else:
            # host-only translation units: plain C++ against the HIP runtime headers
            cmd = [HIPCC, *COMMON, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", "-o", obj, path]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(obj)
    # the translation units compile independently (the RHS instantiations dominate)
    workers = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    shutil.move(tmp, LIB)
    # the command-line driver: a plain C++ consumer of the C ABI, linked against libdmfe.so
    drv_tmp = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(HERE, '..', 'include')}", "-o", drv_tmp,
          os.path.join(CSRC, "dmfe_driver.cpp"), f"-L{HERE}", "-ldmfe", "-Wl,-rpath,$ORIGIN"])
    shutil.move(drv_tmp, DRIVER)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
