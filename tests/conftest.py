import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libsarf.so")


def _gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    """A libsarf context on device 0 (GPU tests only)."""
    if not _gpu_available():
        pytest.skip("no GPU")
    from uncc_sar_focusing_amd.api import Context

    c = Context(0)
    yield c
    c.close()
