import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: CPU runs of many minutes (sanitizer builds); set CARE_SLOW_TESTS=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("CARE_SLOW_TESTS") == "1":
        return
    skip = pytest.mark.skip(reason="slow (sanitizer build of the emulated kernels); set CARE_SLOW_TESTS=1 to run")
    for item in items:
        if "slow" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def root():
    return ROOT
