import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built libjubik_hip.so")


@pytest.fixture(scope="session")
def ja():
    import __graft_entry__ as ge
    return ge.load_package()
