"""
``render()`` on the fused engine with steps queued: the queued steps' last
launch writes the image from the colours its tiles hold in LDS (the render
epilogue), and no separate render pass reads the colours back.  The images
equal the plain engine's (``xentica/core/renderers.py:220-270``,
``base.py:730-753``) for the reference's render golden
(``tests/core/test_base.py:55-70``: crc32(img / 255) = 1955702083), for
colour effects with and without ``SMOOTH_FACTOR``, zooms, pans across the
torus edge and viewports larger than the field.
"""
import binascii

import numpy as np
import pytest

from examples.game_of_life import (
    GameOfLife, GameOfLifeColor, GOLExperiment, GOLExperiment2,
    GOLExperimentColor)
from examples.shifting_sands import ShiftingSands, ShiftingSandsExperiment
from examples.noisetv import NoiseTV, NoiseTVExperiment
from examples.evolife import EvoLife, CrossbreedingExperiment
from xentica_amd.runtime.fused import FusedEngine

pytestmark = pytest.mark.gpu


def view(exp, engine, zoom, pos):
    return type("%s%sZ%d" % (exp.__name__, engine.title(), zoom), (exp, ),
                {"engine": engine, "zoom": zoom, "pos": list(pos)})


def run(model, steps):
    for _ in range(steps):
        model.step()
    return model


@pytest.mark.parametrize("kwargs", [{}, {"fold": False}, {"split": False}])
def test_render_golden_through_the_epilogue(kwargs):
    """1000 steps at depth 7 leave 6 queued: render() runs them with the
    epilogue and the image hashes to the reference's golden."""
    model = GameOfLife(view(GOLExperiment, "fused", 1, (0, 0)))
    model._engine = FusedEngine(model, tile=(64, 16), depth=7, **kwargs)
    model.set_viewport(GOLExperiment.size)
    run(model, 1000)
    assert model._engine._pending == 6
    image = model.render()
    assert model._engine.epilogue_frames == 1
    assert binascii.crc32(image / 255) == 1955702083
    # rendering again (nothing queued) goes through the render kernel
    assert np.array_equal(model.render(), image)
    assert model._engine.epilogue_frames == 1


MODELS = [
    (GameOfLife, GOLExperiment2),
    (GameOfLifeColor, GOLExperimentColor),
    (ShiftingSands, ShiftingSandsExperiment),
    (NoiseTV, NoiseTVExperiment),
    (EvoLife, CrossbreedingExperiment),
]
VIEWS = [
    (1, (0, 0), (640, 360)),
    (3, (17, 5), (200, 120)),
    (2, (630, 350), (333, 77)),
    (1, (600, 300), (1000, 500)),   # wider and taller than the field
    (4, (5, 355), (64, 48)),
]


@pytest.mark.parametrize("cls,exp", MODELS,
                         ids=[cls.__name__ for cls, _ in MODELS])
@pytest.mark.parametrize("zoom,pos,viewport", VIEWS,
                         ids=["z%d_%dx%d" % (z, *v) for z, _, v in VIEWS])
def test_epilogue_frames_vs_plain(cls, exp, zoom, pos, viewport):
    plain = cls(view(exp, "plain", zoom, pos))
    fused = cls(view(exp, "fused", zoom, pos))
    assert isinstance(fused._engine, FusedEngine)
    for model in (plain, fused):
        model.set_viewport(viewport)
    for frame, steps in enumerate((5, 3, 1)):
        run(plain, steps)
        run(fused, steps)
        queued = fused._engine._pending > 0
        before = fused._engine.epilogue_frames
        assert np.array_equal(fused.render(), plain.render()), frame
        assert fused._engine.epilogue_frames == before + queued
    assert fused._engine.epilogue_frames > 0
    assert np.array_equal(fused.gpu.arrays.colors.get(),
                          plain.gpu.arrays.colors.get())


def test_step_render_loop_vs_plain():
    """A viewer's loop, one step per frame: every frame is one 1-step
    launch with the epilogue."""
    plain = GameOfLife(view(GOLExperiment2, "plain", 2, (100, 50)))
    fused = GameOfLife(view(GOLExperiment2, "fused", 2, (100, 50)))
    for model in (plain, fused):
        model.set_viewport((320, 180))
    for frame in range(24):
        plain.step()
        fused.step()
        assert np.array_equal(fused.render(), plain.render()), frame
    assert fused._engine.epilogue_frames == 24
