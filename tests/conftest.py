import os
import sys

# Host LAPACK/BLAS from MKL runs sequentially in every library loaded by the
# tests, so that the reference oracle (LLVM OpenMP) and MKL never bring two
# OpenMP runtimes into one process.
os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402
import pytest  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def small_case():
    """The 1850 x 712 matrix with 8636 uniform draws (duplicates summed),
    N(0,1) values, and its seeded N(0,1) start vector."""
    from propack_amd import synth
    A = synth.random_uniform()
    u0 = synth.start_vector(A.m, 1)
    D = A.todense()
    s = np.linalg.svd(D, compute_uv=False)
    return A, u0, D, s
