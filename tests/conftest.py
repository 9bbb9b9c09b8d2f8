import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available() and torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_ctx():
    if not _gpu_available():
        pytest.skip("no GPU")
    from rayt_amd import api
    ctx = api.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def tmp_scene_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("scenes")
