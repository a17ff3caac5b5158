import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle.pyoracle import ORACLE_LIB

    if not ORACLE_LIB.exists():
        pytest.fail("oracle/liboracle.so missing: run `make -C oracle`")
    return ORACLE_LIB


def rel_maxerr(a, b):
    """max|a-b| / max|b| (max-norm relative error, robust to near-zero entries)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    if scale == 0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / scale)
