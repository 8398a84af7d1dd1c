import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def rel_err(a, b, floor=0.0):
    """Per-row relative error |a-b| / max(|b|, floor)."""
    a = np.asarray(a).reshape(-1, 3)
    b = np.asarray(b).reshape(-1, 3)
    den = np.maximum(np.linalg.norm(b, axis=1), floor)
    return np.linalg.norm(a - b, axis=1) / den


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    return oracle
