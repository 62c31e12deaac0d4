import os
import sys

# The reference build's scalar quadratures are OpenMP reductions (convolution.cpp:13):
# the oracle and the goldens follow the one-thread order.  libgomp reads this once.
os.environ["OMP_NUM_THREADS"] = "1"

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
if os.path.join(ROOT, "tests") not in sys.path:
    sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "ref: needs the reference build under oracle/_ref (CPU)")
