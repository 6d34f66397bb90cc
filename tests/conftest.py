import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def gpu_available() -> bool:
    """True when a HIP device is visible (checked without touching it from
    Python: libgomc_hip.so's box creation reports GOMC_ERR_NODEV otherwise)."""
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    return pyoracle.lib()


def rel_close(a, b, rtol, scale=None):
    """|a-b| <= rtol * max(|b|, scale, 1e-300) elementwise."""
    import numpy as np
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    s = np.maximum(np.abs(b), 1e-300 if scale is None else scale)
    return np.all(np.abs(a - b) <= rtol * s)
