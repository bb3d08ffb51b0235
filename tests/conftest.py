"""Shared pytest setup: the `gpu` marker and paths to the oracle and the library."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "oracle"), os.path.join(REPO, "irbc-code_amd"),
          os.path.join(REPO, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def repo():
    return REPO


@pytest.fixture(scope="session")
def oracle_lib():
    import pyoracle
    pyoracle.lib()
    return pyoracle
