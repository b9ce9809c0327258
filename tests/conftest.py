import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "propag-hidalgo_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (HIP)")


def pytest_collection_modifyitems(config, items):
    # Print the case name before each GPU test runs (flushed), so a run that
    # ends the box still says what was running.
    pass


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.build(ref=False)
    return oracle


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "firelib_fp64.npz"))


@pytest.fixture(autouse=True)
def _announce(request):
    if request.node.get_closest_marker("gpu"):
        print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
