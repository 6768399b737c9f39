import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")
    config.addinivalue_line("markers", "slow: long-running case")


def _lib_path():
    return os.path.join(ROOT, "zquatev_amd", "libzquatev_amd.so")


@pytest.fixture(scope="session", autouse=True)
def _built_library():
    """Tests load what is built; build only when nothing is."""
    if not os.path.exists(_lib_path()):
        import __graft_entry__
        __graft_entry__.build()
    yield


@pytest.fixture(scope="session")
def gpu_ctx():
    import zquatev_amd as zq
    ctx = zq.Context(0)
    yield ctx
    ctx.close()
