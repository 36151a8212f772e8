import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def pytest_collection_modifyitems(config, items):
    # GPU tests name their case before running so a device fault is attributable
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(pytest.mark.timeout(300))


@pytest.fixture(autouse=True)
def _announce(request):
    if "gpu" in request.node.keywords:
        print(f"\n[case] {request.node.nodeid}", flush=True)
    yield
