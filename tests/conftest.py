"""Test configuration: import paths, the ``gpu`` marker, shared fixtures.

``-m "not gpu"`` runs here without a GPU (oracle against golden vectors and closed forms,
host logic, library symbol exports).  ``-m gpu`` runs the HIP path against the oracle on a
ROCm device.
"""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "metal-singlecell_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libmsc_hip.so")


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no ROCm GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def c1_data():
    from metalsinglecell import synthetic
    return synthetic.gaussian_mixture_isotropic()
