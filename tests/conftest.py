import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running parity case")


def pytest_collection_modifyitems(config, items):
    # fixed order: oracle/host tests first, GPU parity after
    pass


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library with a visible device; fails (not skips) when it is missing."""
    from scalingelections_amd import _lib, api

    _lib.lib()
    if api.device_count() < 1:
        pytest.fail("no HIP device visible to the library")
    return _lib
