import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libgale_stokes.so")


@pytest.fixture(scope="session")
def gale():
    """The HIP library; GPU tests fail (not skip) when it or the device is missing."""
    import gale_amd
    n = gale_amd.device_count()
    assert n > 0, "no HIP device visible: GPU tests need an MI355X"
    return gale_amd


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


def load_csr(name):
    import scipy.sparse as sp
    z = np.load(os.path.join(GOLDEN, name))
    return sp.csr_matrix((z["data"], z["indices"], z["indptr"]), shape=tuple(z["shape"]))


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


try:  # import once at collection so the first GPU test's timeout does not include it
    import torch  # noqa: F401
except Exception:  # pragma: no cover
    pass
