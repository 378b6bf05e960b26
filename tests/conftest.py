import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def engine_lib():
    import gpu_2dising_amd as G

    G.lib()
    return G


@pytest.fixture(scope="session")
def gpu(engine_lib):
    """Initialises the device once; GPU tests fail loudly when none is visible."""
    ncu = engine_lib.init_device(-1)
    return ncu
