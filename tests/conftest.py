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
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    rp = np.asarray(rp, dtype=np.int64)
    starts = rp[:-1][np.diff(rp) > 0]  # non-empty rows: each segment ends where the next begins
    if len(starts) == 0:
        return 0.0
    scale = np.maximum.reduceat(np.abs(b), starts)
    diff = np.maximum.reduceat(np.abs(a - b), starts)
    ok = scale > 0
    return float((diff[ok] / scale[ok]).max()) if ok.any() else 0.0
