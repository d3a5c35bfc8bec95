import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built libfdlbm.so")


def pytest_collection_modifyitems(config, items):
    # GPU tests print what they run so a failing GPU step can be named from
    # the log alone.
    pass


@pytest.fixture(autouse=True)
def _name_the_case(request):
    if request.node.get_closest_marker("gpu") is not None:
        print("\n[gpu-case] %s" % request.node.nodeid, flush=True)
    yield
