import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def fl():
    """The product library (fails loudly when it is not built or no GPU)."""
    from fluidloom_amd import _lib
    L = _lib.lib()
    n = _lib.device_count()
    assert n >= 1, "no GPU visible"
    _lib.check(L.fl_init(0))
    return _lib


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
