import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"
REFBIN = ROOT / "oracle" / "_ref"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libradgpu.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def refbin():
    if not (REFBIN / "rcontrib").exists():
        pytest.skip("reference tools not built (make -C oracle ref)")
    return REFBIN


@pytest.fixture(scope="session")
def gpu():
    from pyradiance_amd import _lib
    return _lib
