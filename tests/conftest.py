import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libgoss_hip.so")


@pytest.fixture(autouse=True)
def _announce(request):
    # GPU steps print their case before running so that a failing box still
    # names it in the captured output.
    if request.node.get_closest_marker("gpu"):
        print(f"\n[case] {request.node.nodeid}", flush=True)
    yield


def golden(name):
    return os.path.join(GOLDEN, name)
