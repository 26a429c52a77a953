import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libopenblas_amd.so")


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "level3_golden.npz"))


@pytest.fixture(scope="session")
def oba():
    """The product library (GPU tests only): fails loudly if not built."""
    import openblas_amd
    L = openblas_amd.lib()
    if openblas_amd.lib().openblas_amd_device_count() < 1:
        pytest.fail("no HIP device visible to libopenblas_amd")
    return openblas_amd


@pytest.fixture(scope="session")
def dsgesv_fallback_cases():
    """Inputs that take dsgesv.f's two other fallback exits: an entry of B
    or A beyond FLT_MAX (DLAG2S overflow, ITER = -2, dsgesv.f:291-305)
    and an A whose single-precision copy has an all-zero column (entries
    below FLT_TRUE_MIN demote to 0 without an overflow flag; SGETRF then
    returns INFO > 0, ITER = -3, dsgesv.f:310-314)."""
    rng = np.random.default_rng(23)
    n = 40
    base = rng.standard_normal((n, n)) + 10 * np.eye(n)
    rhs = rng.standard_normal((n, 2))
    cases = {}
    B = rhs.copy()
    B[3, 1] = 1e39
    cases["B_overflow"] = (base.copy(), B, -2)
    A = base.copy()
    A[5, 7] = -1e39
    A[7, 7] = 1e40
    cases["A_overflow"] = (A, rhs.copy(), -2)
    A = base.copy()
    A[:, 4] *= 1e-50
    cases["A_single_singular"] = (A, rhs.copy(), -3)
    return cases
