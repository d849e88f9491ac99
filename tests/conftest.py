import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def ctx():
    import afepack_amd as afe
    # objects created from the context hold a reference to it, so it is
    # destroyed after them by reference counting
    yield afe.Context(0)


def make_mesh(kind, n, jitter=0.0, seed=0):
    import afepack_amd as afe
    if kind == afe.P1_TET:
        return afe.mesh_cube(n, jitter, seed)
    return afe.mesh_square(n, jitter, seed)


def rel_err_rows(a, b, rp):
    """max over rows of |a-b| / max|b in row| (entries that cancel to ~0 are
    measured against the row scale)"""
    worst = 0.0
    for i in range(len(rp) - 1):
        s = slice(rp[i], rp[i + 1])
        scale = np.abs(b[s]).max()
        if scale > 0:
            worst = max(worst, np.abs(a[s] - b[s]).max() / scale)
    return worst
