"""
render() with the reference's int32 -> uint8 cast done on the device
(``xentica/core/base.py:753``: ``img.get().astype(numpy.uint8)``).
"""
import numpy as np
import pytest

from examples.game_of_life import GameOfLifeColor, GOLExperimentColor
from xentica_amd.runtime import hip, kernels

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("count", [1, 3, 4, 1001, 3 * 4096 + 2])
def test_low_bytes_matches_numpy_cast(count):
    rng = np.random.default_rng(count)
    values = rng.integers(-2 ** 31, 2 ** 31, count, dtype=np.int64)
    values = values.astype(np.int32)
    values[:min(count, 6)] = [0, 255, 256, -1, -256, 2 ** 31 - 1][:min(count, 6)]
    src, dst = hip.Buffer(4 * count), hip.Buffer(count)
    src.upload(values)
    kernels.low_bytes(src.ptr, dst.ptr, count)
    got = np.empty(count, dtype=np.uint8)
    dst.download(got)
    assert np.array_equal(got, values.astype(np.uint8))


@pytest.mark.parametrize("engine", ["plain", "fused"])
def test_render_equals_cast_of_int32_image(engine):
    exp = type("Render" + engine, (GOLExperimentColor, ), {"engine": engine})
    model = GameOfLifeColor(exp)
    model.set_viewport((200, 120))
    for _ in range(5):
        model.step()
    frame = model.render()
    assert frame.dtype == np.uint8 and frame.shape == (200 * 120 * 3, )
    assert np.array_equal(frame, model.gpu.arrays.img.get().astype(np.uint8))
    model.set_viewport((64, 32))  # a smaller image reuses the staging buffer
    assert np.array_equal(model.render(),
                          model.gpu.arrays.img.get().astype(np.uint8))
