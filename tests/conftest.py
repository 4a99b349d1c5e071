"""Shared fixtures.  Registers the ``gpu`` marker and makes the repository
root importable so tests can load the package and the oracle."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def me():
    """The product package (motion-estimation_amd/), loaded by path."""
    import __graft_entry__ as ge

    return ge.load_package()


def _has_gpu():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


HAS_GPU = _has_gpu()
