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
