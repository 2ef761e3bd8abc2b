import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _pkg():
    return importlib.import_module("kynema-sgf_amd")


@pytest.fixture(scope="session")
def ksgf():
    return _pkg()


@pytest.fixture(scope="session")
def workloads():
    return importlib.import_module("kynema-sgf_amd.workloads")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.build()
    return O


@pytest.fixture(scope="session")
def gpu(ksgf):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    ksgf.load()
    return torch
