import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhcgpu.so on the device)")


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work so a failing box can be traced to it
    if request.node.get_closest_marker("gpu"):
        print("\n[gpu-test] %s" % request.node.nodeid, flush=True)
    yield
