"""Shared pytest set-up: the `gpu` marker, import paths, and the oracle
library (test infrastructure) built on first use."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


def rel(a, b):
    """max |a - b| / max |b| (absolute when b is all zero)."""
    import numpy as np
    a, b = np.asarray(a), np.asarray(b)
    s = np.abs(b).max() if b.size else 0.0
    d = np.abs(a - b).max() if a.size else 0.0
    return d / s if s > 0 else d
