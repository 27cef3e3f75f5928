"""The C-ABI library loads without a GPU and exports every symbol include/msc_hip.h declares;
argument validation and workspace queries run on the host (no GPU calls)."""

import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "msc_hip.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(msc_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_path():
    names = _declared()
    for want in ("msc_knn_exact", "msc_fuzzy_connectivities", "msc_leiden", "msc_spatial_knn",
                 "msc_moran", "msc_last_error"):
        assert want in names


def test_library_exports_every_declared_symbol():
    from metalsinglecell import _hip
    lib = _hip.load_library()
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(_declared()) <= set(_hip.SIGNATURES)
    assert lib.msc_version() >= 100


def test_validation_errors_without_gpu():
    from metalsinglecell import _hip
    lib = _hip.load_library()
    nb = ctypes.c_size_t(0)
    assert lib.msc_knn_workspace_size(1000, 50, 30, 1, ctypes.byref(nb)) == 0 and nb.value > 0
    assert lib.msc_knn_workspace_size(10, 50, 30, 1, ctypes.byref(nb)) != 0
    assert b"exceeds" in lib.msc_last_error()
    # wide shapes (k > 64, d beyond the MFMA K padding) take the LDS-list path: token workspace
    assert lib.msc_knn_workspace_size(1000, 50, 65, 0, ctypes.byref(nb)) == 0 and nb.value == 256
    assert lib.msc_knn_workspace_size(1000, 200, 15, 1, ctypes.byref(nb)) == 0 and nb.value == 256
    assert lib.msc_knn_workspace_size(1000, 50, 513, 1, ctypes.byref(nb)) != 0
    assert b"k=513" in lib.msc_last_error()
    assert lib.msc_fuzzy_workspace_size(1000, 15, ctypes.byref(nb)) == 0
    assert lib.msc_fuzzy_workspace_size(1000, 1025, ctypes.byref(nb)) != 0
    assert lib.msc_leiden_workspace_size(1000, 30000, ctypes.byref(nb)) == 0
    assert lib.msc_spatial_knn_workspace_size(1000, 2, 6, ctypes.byref(nb)) == 0
    assert lib.msc_spatial_knn_workspace_size(1000, 4, 6, ctypes.byref(nb)) == 0 and nb.value == 256
    assert lib.msc_spatial_knn_workspace_size(1000, 2, 40, ctypes.byref(nb)) == 0 and nb.value == 256
    assert lib.msc_spatial_knn_workspace_size(1000, 0, 6, ctypes.byref(nb)) != 0
    assert lib.msc_moran_workspace_size(1000, 500, 6000, 100, ctypes.byref(nb)) == 0


def test_gpu_calls_raise_without_a_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import numpy as np
    from metalsinglecell import _hip
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _hip.knn_exact(np.zeros((10, 3), np.float32), 3)
