import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as GE  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X / HIP device")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def dim3():
    return GE.load_package()


@pytest.fixture(scope="session")
def ctx(dim3):
    c = dim3.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def torch():
    import torch as _torch

    return _torch
