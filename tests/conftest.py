import os
import pathlib
import sys

import pytest

REPO = pathlib.Path(__file__).resolve().parent.parent
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _ensure_built():
    """Load what is built; build only when a library is missing (CPU container)."""
    from fv_140823_amd import build as B
    if not B.LIB.exists():
        B.build_library()
    if not B.ORACLE_LIB.exists():
        B.build_oracle()


_ensure_built()


@pytest.fixture(scope="session")
def has_gpu():
    import torch
    return torch.cuda.is_available()
