"""Shared test setup: import paths, the `gpu` marker and helpers."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent

# torch (when importable) brings up its HIP runtime before any test module
# loads libvtp.so, which then binds to that same runtime; the other order
# leaves torch reporting no GPU, and the torch-tensor tests need both.
try:
    import torch
    torch.cuda.is_available()
except Exception:
    pass

sys.path.insert(0, str(ROOT / "vt-physics_amd"))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvtp.so on cuda:0)")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def rel_err(a, b) -> float:
    """max |a - b| / max |b| (norm-wise relative error of a field)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.max(np.abs(b))
    if scale == 0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / scale)


@pytest.fixture(scope="session")
def lib():
    import vtp
    return vtp.load_library()
