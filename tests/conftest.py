import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    # the oracle is test infrastructure: build it if this checkout has not yet
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so", "liboracle_fast.so"],
                       check=True, capture_output=True)


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def pdm():
    from engine_amd import paradmm
    paradmm.load()
    return paradmm


def rel_err(a, b):
    """max |a - b| / max |b| (norm-wise relative error of one snapshot)."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
