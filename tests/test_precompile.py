"""
Parallel precompilation into the code-object cache (hiprtc runs without a
GPU): every source lands in the cache, equal to a serial compile.
"""
import os

from examples.workloads import LifeLikeRule, batch_experiments
from xentica_amd.runtime import hip


def _code_objects(path):
    # the cache folder also holds the host backend's libraries (``hostc/``),
    # which building a model's rule probe may create
    return [name for name in os.listdir(path) if name.endswith(".co")]


def test_precompile_fills_the_cache(tmp_path, monkeypatch):
    monkeypatch.setenv("XENTICA_CACHE_DIR", str(tmp_path))
    sources = [LifeLikeRule(exp).cuda_source
               for exp in batch_experiments(count=5, side_min=40,
                                            side_max=90)]
    options = hip.compile_options("off")
    unique = len(set(sources))
    assert hip.precompile(sources + sources[:2], options, workers=3) == unique
    assert len(_code_objects(tmp_path)) == unique
    # a second call finds everything cached
    assert hip.precompile(sources, options) == 0
    cached = hip.compile_source(sources[3], "again.hip", options)
    os.remove(os.path.join(tmp_path, _code_objects(tmp_path)[0]))
    monkeypatch.setenv("XENTICA_CACHE_DIR", str(tmp_path / "fresh"))
    assert hip.compile_source(sources[3], "serial.hip", options) == cached
