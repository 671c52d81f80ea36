import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref (the flang-built reference)")


@pytest.fixture(scope="session")
def bohr():
    return 0.52917721067  # main.f90:59
