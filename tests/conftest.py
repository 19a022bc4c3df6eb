import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for all GPU tests (a single process holds the GPU)."""
    from skyrendering_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def blue_noise_64():
    return np.fromfile(os.path.join(GOLDEN, "blue_noise_64x64.u16"), dtype="<u2").reshape(64, 64)


def rel_err(gpu, ref, floor_frac=1e-6):
    """Per-element |a - b| / max(|b|, floor_frac * max|b|); returns (max, rms, argmax)."""
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.maximum(np.abs(ref), floor_frac * np.max(np.abs(ref)))
    scale = np.where(scale == 0, 1.0, scale)
    e = np.abs(gpu - ref) / scale
    return float(e.max()), float(np.sqrt(np.mean(e * e))), np.unravel_index(int(np.argmax(e)), e.shape)
