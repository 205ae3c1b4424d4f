import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    from oracle import Oracle

    return Oracle("f64")


@pytest.fixture(scope="session")
def orc32():
    from oracle import Oracle

    return Oracle("f32")


@pytest.fixture(scope="session")
def mm():
    return importlib.import_module("1_amd")


@pytest.fixture(scope="session")
def lib(mm):
    """The HIP library; GPU tests fail loudly if it is missing."""
    return mm.load()
