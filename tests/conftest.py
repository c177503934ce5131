"""pytest configuration: the `gpu` marker and shared fixtures.

`python -m pytest tests -m "not gpu"` runs on a machine without a GPU: the
oracle against the reference's golden fixtures, the host logic, and the C-ABI
(library load, exported symbols, argument checks that return before any GPU
work).  `-m gpu` runs the parity tests proper through the C-ABI on an MI355X.
"""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")


@pytest.fixture(scope="session")
def oracle():
    import _oracle
    return _oracle


@pytest.fixture(scope="session")
def golden():
    import _golden
    return _golden
