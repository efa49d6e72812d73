import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def lib():
    from cudadsmc_amd import _abi
    return _abi.load_library()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.load()
    return oracle
