import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm) GPU and the built HIP library")


def relmax(a, b):
    """max|a - b| / max|b| (tests/gates/mhs/_common.py:60-66)."""
    import numpy as np
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    ref = np.max(np.abs(b))
    diff = np.max(np.abs(a - b))
    return diff / ref if ref > 1e-300 else diff


@pytest.fixture(scope="session")
def small():
    import numpy as np
    return dict(np.load(GOLDEN / "small.npz"))
