import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def golden_sm160():
    import numpy as np
    return dict(np.load(os.path.join(GOLDEN, "sm160_golden.npz")))


@pytest.fixture(scope="session")
def golden_hex8():
    import numpy as np
    return dict(np.load(os.path.join(GOLDEN, "hex8_golden.npz")))


def soa(Fs):
    """[n,3,3] (row, col) -> SoA [9,n], column-major per point."""
    import numpy as np
    return np.ascontiguousarray(Fs.transpose(0, 2, 1).reshape(-1, 9).T)


def unsoa(Fsoa):
    """SoA [9,n] column-major -> [n,3,3] (row, col)."""
    import numpy as np
    return np.ascontiguousarray(Fsoa.T.reshape(-1, 3, 3).transpose(0, 2, 1))
