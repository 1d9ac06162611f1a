"""Test configuration: the ``gpu`` marker and import paths."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) and libxentica_hip.so")


@pytest.fixture(autouse=True)
def _cache_dir(tmp_path_factory, monkeypatch):
    """Keep compiled code objects in one per-session temporary folder."""
    cache = tmp_path_factory.getbasetemp() / "xentica_cache"
    cache.mkdir(exist_ok=True)
    monkeypatch.setenv("XENTICA_CACHE_DIR", str(cache))


@pytest.fixture(autouse=True)
def _gpu_progress(request):
    """Name each GPU test on stderr before it touches the device."""
    import time
    if request.node.get_closest_marker("gpu") is None:
        yield
        return
    start = time.time()
    sys.__stderr__.write("[gpu-test start] %s\n" % request.node.nodeid)
    sys.__stderr__.flush()
    yield
    sys.__stderr__.write("[gpu-test end] %s %.1fs\n" % (
        request.node.nodeid, time.time() - start))
    sys.__stderr__.flush()
