import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE_DIR = os.path.join(ROOT, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def pkg():
    return importlib.import_module("4dzb_amd")


@pytest.fixture(scope="session")
def wl(pkg):
    return pkg.workloads


@pytest.fixture(scope="session")
def fd(pkg):
    return pkg.fd4d


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    return o
