import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle_ctypes
    oracle_ctypes.load("det")
    return oracle_ctypes


@pytest.fixture(scope="session")
def qp_ctx():
    from quetzalpic_amd import capi
    ctx = capi.Context(0)
    yield ctx
    ctx.close()


def rel_err(a, b):
    """(max |a-b| / max |b|, ||a-b||_2 / ||b||_2)."""
    a = np.asarray(a)
    b = np.asarray(b)
    d = a - b
    mb = np.abs(b).max()
    nb = np.linalg.norm(b)
    return (np.abs(d).max() / mb if mb > 0 else np.abs(d).max(),
            np.linalg.norm(d) / nb if nb > 0 else np.linalg.norm(d))
