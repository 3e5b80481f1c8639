import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpu-svdag-editing_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


def _log(msg):
    print(msg, flush=True)


@pytest.fixture(autouse=True)
def _announce(request):
    # Name the case before any GPU step so a box that dies still tells us which.
    if request.node.get_closest_marker("gpu"):
        _log(f"\n[gpu-case] {request.node.nodeid}")
    yield
