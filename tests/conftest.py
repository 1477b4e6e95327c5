import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libswinmd_hip.so)")


@pytest.fixture(scope="session")
def oracle_mod():
    from tests.helpers import ensure_oracle
    return ensure_oracle()
