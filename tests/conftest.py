import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Device check for gpu-marked tests: skip only when no GPU is present; a
    missing library on a GPU machine is an error, not a skip."""
    if not gpu_available():
        pytest.skip("no GPU")
    from physgrad_amd import _lib
    _lib.load()
    return True


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    O.lib()
    return O
