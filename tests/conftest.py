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
