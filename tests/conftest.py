import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "fringe-sgc_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def fsgc_mod():
    import fsgc
    return fsgc


@pytest.fixture(scope="session")
def gpu(fsgc_mod):
    """The library initialised on device 0. A GPU test fails (does not skip) when the
    device cannot be initialised: the HIP path is the product and has no fallback."""
    fsgc_mod.init(0)
    return fsgc_mod
