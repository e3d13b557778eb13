"""Shared test set-up.

``-m "not gpu"`` runs the oracle pins, host logic and ABI checks (no GPU).
``-m gpu`` runs the parity tests of the HIP path against the oracle on an
MI355X; they fail (not skip) when the built library or the GPU is missing.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "warpflow-les_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    """Selects GPU 0 through the HIP library; raises if either is missing."""
    from warpflow import _hip
    _hip.ensure_device(0)
    return _hip


def rel_err(a, b, ref=None):
    """max|a-b| / max|ref| (ref defaults to b); 0 when both vanish."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b if ref is None else np.asarray(ref)).max(initial=0.0)
    d = np.abs(a - b).max(initial=0.0)
    if scale == 0.0:
        return d
    return d / scale
