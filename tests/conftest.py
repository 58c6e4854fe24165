import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as _entry  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def pkg():
    return _entry.load_package()


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (checker)."""
    return _entry.load_oracle()


@pytest.fixture(scope="session")
def T(pkg):
    return pkg._types


def rel_l2(a, b):
    import numpy as np

    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (nb if nb > 0 else 1.0)
