"""
Lifetimes of runtime objects around stream capture (CPU, fake library).

A module must be released by reference counting, not by the cyclic
collector, and the collector must not run while a stream is captured:
either would run ``__del__`` (module unload, ``hipFree``) on the
capturing thread, which invalidates the capture.
"""
import gc
import weakref

import pytest

from xentica_amd.runtime import hip


class _FakeLib:
    def __init__(self):
        self.unloaded = 0

    def xen_module_function(self, module, name, out):
        return 0

    def xen_module_unload(self, handle):
        self.unloaded += 1

    def xen_graph_begin(self, stream):
        return 0

    def xen_graph_end(self, stream, out):
        return 0


@pytest.fixture
def fake_lib(monkeypatch):
    lib = _FakeLib()
    monkeypatch.setattr(hip, "library", lambda: lib)
    monkeypatch.setattr(hip, "_LIB", lib)
    return lib


def _module():
    module = hip.Module.__new__(hip.Module)
    module.handle = hip._vp(1)
    module._handles = {}
    return module


def test_module_freed_without_the_cyclic_collector(fake_lib):
    was = gc.isenabled()
    gc.disable()
    try:
        module = _module()
        func = module.get_function("fused_bsca")
        assert func.module is module
        ref = weakref.ref(module)
        del module, func
        assert ref() is None
        assert fake_lib.unloaded == 1
    finally:
        if was:
            gc.enable()


class _Stream:
    handle = None


def test_collector_paused_during_capture(fake_lib):
    assert gc.isenabled()
    hip.Graph.begin(_Stream())
    assert not gc.isenabled()
    graph = hip.Graph.end(_Stream())
    assert gc.isenabled()
    graph.handle = None
