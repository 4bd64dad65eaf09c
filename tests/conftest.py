import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a visible MI355X (gfx950) GPU")


def gpu_visible() -> bool:
    try:
        import acceles_amd as acc
        return acc.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def acc():
    import acceles_amd
    acceles_amd.lib()
    return acceles_amd


@pytest.fixture(scope="session")
def gpu(acc):
    """Skips only when no GPU is visible; on a GPU box a missing library is an error."""
    if acc.device_count() == 0:
        pytest.skip("no GPU visible")
    return acc
