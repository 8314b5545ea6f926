import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libptg.so on cuda:0)")


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def engine():
    from ptarmigan_amd import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    return lambda name: np.load(os.path.join(here, name + ".npz"))


@pytest.fixture(scope="session")
def golden_tables():
    """The packed table blob the library and the oracle load (ptarmigan_amd/tables)."""
    from ptarmigan_amd import tables
    return tables.load()
