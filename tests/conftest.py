import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "qclnegf.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def rel(x, y):
    """Norm-relative difference, the reference's isapprox(x, y; rtol) measure."""
    x = np.asarray(x)
    y = np.asarray(y)
    d = np.linalg.norm((x - y).ravel())
    n = max(np.linalg.norm(x.ravel()), np.linalg.norm(y.ravel()))
    return d / n if n > 0 else d


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.build()
    return oracle
