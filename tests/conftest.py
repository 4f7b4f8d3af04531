import os
import sys

# The oracle's SuperLU factorizations call BLAS on small dense blocks.  A multi-threaded
# BLAS pool spin-waits for its workers; with another busy process on the host those
# waits stretched a 5 s test past two minutes.  One BLAS thread per test process.
for _v in ("OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ.setdefault(_v, "1")

import pytest  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


def golden(name: str) -> str:
    return os.path.join(GOLDEN, name)
