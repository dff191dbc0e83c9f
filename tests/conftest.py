import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def rx():
    """The product package; builds librxhip.so only if nothing is built."""
    from renderx_amd import _lib, api

    if not os.path.exists(_lib.LIB_PATH):
        from renderx_amd import build

        build.build()
    return api


@pytest.fixture(scope="session")
def gpu_ctx(rx):
    return rx.Context(0)
