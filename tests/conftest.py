import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def engine():
    """One engine context for the whole GPU session (one process on the card)."""
    from ivpsnr_amd import Engine
    from ivpsnr_amd.capi import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.fail("libqmiv.so is not built; run __graft_entry__.build() first")
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        pytest.fail("oracle library is not built; run __graft_entry__.build() first")
    return pyoracle
