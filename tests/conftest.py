import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu_ctx():
    from rayt_amd import api
    try:
        ctx = api.Context(0)
    except api.RaytError as e:
        pytest.skip(f"no GPU: {e}")
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def tmp_scene_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("scenes")
