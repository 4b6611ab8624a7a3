import importlib
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def cpmp():
    """The product package (fails loudly if libcpmp.so is not built)."""
    return importlib.import_module("3dcpmpmodel_amd")


@pytest.fixture(scope="session")
def orc():
    import oracle_lib
    return oracle_lib


@pytest.fixture(scope="session")
def cocurve():
    with open(os.path.join(ROOT, "tests", "golden", "cocurve.json")) as fh:
        return json.load(fh)["tables"]


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
