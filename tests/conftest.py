import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def behler():
    from libdescriptor_amd.workloads import behler_si_params
    return behler_si_params()


def rel_err(a, b, floor=0.0):
    """max |a-b| / max(|b|_inf, floor): normwise relative error."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.abs(b).max() if b.size else 0.0, floor)
    return 0.0 if den == 0 else float(np.abs(a - b).max() / den)


def elem_rel_err(a, b, abs_floor):
    """max over entries of |a-b| / max(|b|, abs_floor)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float((np.abs(a - b) / np.maximum(np.abs(b), abs_floor)).max()) if b.size else 0.0


@pytest.fixture(scope="session")
def c1():
    return golden("c1_ref.npz")
