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
    try:  # tests hand torch device arrays to the library: torch's HIP runtime starts first
        import torch  # noqa: F401
    except ImportError:
        pass
    from plantfem_amd import _lib as L
    return L.default_context()
