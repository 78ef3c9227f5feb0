"""Shared fixtures.  Tests marked `gpu` need a HIP device (MI355X); the rest run on CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (runs on the MI355X box)")


@pytest.fixture(scope="session")
def gr():
    import graphite_amd
    graphite_amd.load()
    return graphite_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def normwise(a, b):
    """||a-b|| / ||b|| (Frobenius), 0 when both are zero."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return 0.0 if d == 0 else d / (nb if nb > 0 else 1.0)
