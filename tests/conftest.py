import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C-ABI)")


@pytest.fixture(scope="session")
def ctx():
    from plantfem_amd import _lib as L
    return L.default_context()
