import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    """Relative Frobenius error of a against the reference b."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def align_columns(a, ref):
    """Flip column signs of a so that every column agrees in sign with ref."""
    a = np.array(a, dtype=np.float64, copy=True)
    for j in range(min(a.shape[1], ref.shape[1])):
        if float(a[:, j] @ ref[:, j]) < 0:
            a[:, j] = -a[:, j]
    return a


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle, build
    build()
    return Oracle()


@pytest.fixture(scope="session")
def hip():
    """The HIP library; GPU tests fail loudly (never skip) when it is missing."""
    import fastpls_amd
    fastpls_amd.load_library()
    info = fastpls_amd.device_info()
    return fastpls_amd, info
