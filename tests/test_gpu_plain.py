"""
The plain engine on the GPU against the reference's golden checksums.

Each test restates a test of ``tests/core/test_base.py`` or
``tests/core/test_properties.py`` of the reference with the same model,
experiment, step count and expected CRC-32, and runs it through the
generated HIP kernels (hiprtc, gfx950).  The C1 workload (Life 512x512,
seed 42) is checked against the CPU oracle after every step.
"""
import binascii
import os

import numpy as np
import pytest

from xentica_amd.core.exceptions import XenticaException
from xentica_amd.core.properties import IntegerProperty
from examples.game_of_life import (
    GameOfLife, GameOfLifeStatic, GameOfLifeColor, GameOfLife6D,
    GOLExperiment, GOLExperiment2, GOLExperimentColor,
    LifelikeCA, DiamoebaExperiment,
)
from examples.shifting_sands import ShiftingSands, ShiftingSandsExperiment
from examples.noisetv import NoiseTV, NoiseTVExperiment
from examples.workloads import Life512Experiment

pytestmark = pytest.mark.gpu

NUM_STEPS = 1000


@pytest.fixture(autouse=True)
def _plain_engine(monkeypatch):
    """Every model in this file runs on the reference-structured engine."""
    from xentica_amd.core.base import CellularAutomaton
    monkeypatch.setattr(CellularAutomaton, "engine", "plain")


def run(model, steps):
    for _ in range(steps):
        model.step()
    return model


def cells_crc(model):
    return binascii.crc32(model.gpu.arrays.cells.get()[:model.cells_num])


def test_single_ca():
    """``test_base.py:28-41``."""
    from xentica_amd.runtime.engines import PlainEngine
    model = run(GameOfLife(GOLExperiment), NUM_STEPS)
    assert isinstance(model._engine, PlainEngine)
    assert cells_crc(model) == 2981695958


def test_multiple_ca():
    """``test_base.py:43-53``: two models interleaved in one process."""
    mod1 = GameOfLife(GOLExperiment)
    mod2 = GameOfLife(GOLExperiment)
    for _ in range(NUM_STEPS):
        mod1.step()
        mod2.step()
    assert cells_crc(mod1) == 2981695958
    assert cells_crc(mod2) == 2981695958


def test_pause():
    """``test_base.py:72-90``."""
    model = GameOfLife(GOLExperiment)
    before = cells_crc(model)
    model.step()
    after = cells_crc(model)
    assert before != after
    model.paused = True
    model.step()
    assert cells_crc(model) == after


def test_save_load(tmp_path):
    """``test_base.py:92-104``: half run, save, load into a new model, finish."""
    path = str(tmp_path / "test.ca")
    ca1 = run(GameOfLife(GOLExperiment), NUM_STEPS // 2)
    ca1.save(path)
    ca2 = GameOfLife(GOLExperiment)
    ca2.load(path)
    run(ca2, NUM_STEPS // 2)
    assert cells_crc(ca2) == 2981695958


def test_load_random(tmp_path):
    """``test_base.py:106-118``."""
    path = str(tmp_path / "test.ca")
    ca1 = GameOfLife(GOLExperiment)
    ca1.save(path)
    ca2 = GameOfLife(GOLExperiment)
    ca2.load(path)
    assert ca1.random.standard.randint(1, 1000) == \
        ca2.random.standard.randint(1, 1000)
    assert ca1.random.numpy.randint(1, 1000) == \
        ca2.random.numpy.randint(1, 1000)


def test_static_border():
    """``test_base.py:138-147``."""
    model = run(GameOfLifeStatic(GOLExperiment), NUM_STEPS)
    assert cells_crc(model) == 1098273940


def test_multiple_properties():
    """``test_base.py:149-156``."""
    model = run(GameOfLifeColor(GOLExperimentColor), NUM_STEPS)
    assert cells_crc(model) == 532957133


def test_multidimensional():
    """``test_base.py:158-168``."""
    class GOLExperiment6DLite(GOLExperiment2):
        size = (64, 36, 3, 3, 3, 3)
    model = run(GameOfLife6D(GOLExperiment6DLite), NUM_STEPS)
    assert cells_crc(model) == 2742543959


def test_cell_width():
    """``test_base.py:170-180``: 16-bit state, cast to uint8 before the CRC."""
    class GameOfLifeInt(GameOfLife):
        state = IntegerProperty(max_val=2 ** 16 - 1)
    model = run(GameOfLifeInt(GOLExperiment), NUM_STEPS)
    cells = model.gpu.arrays.cells.get()[:model.cells_num].astype(np.uint8)
    assert binascii.crc32(cells) == 2981695958


def test_nonuniform_interactions():
    """``test_base.py:182-189``."""
    model = run(ShiftingSands(ShiftingSandsExperiment), NUM_STEPS)
    assert cells_crc(model) == 1639048270


def test_unset_viewport():
    """``test_base.py:191-195``."""
    with pytest.raises(XenticaException):
        ShiftingSands(ShiftingSandsExperiment).render()


def test_interactive():
    """``test_base.py:218-225``."""
    model = run(LifelikeCA(DiamoebaExperiment), NUM_STEPS)
    assert cells_crc(model) == 4016906756


def test_random_property():
    """``test_properties.py:64-71``."""
    model = run(NoiseTV(NoiseTVExperiment), 100)
    assert cells_crc(model) == 2773894957


def test_render():
    """``test_base.py:55-70``: crc32(img / 255) of a zoom-1 render."""
    class RenderExperiment(GOLExperiment):
        zoom = 1
    model = GameOfLife(RenderExperiment)
    model.set_viewport(RenderExperiment.size)
    run(model, NUM_STEPS)
    img = model.render()
    assert binascii.crc32(img / 255) == 1955702083


def test_life512_every_step_vs_oracle():
    """Life 512x512 seed 42: hash of every one of 1000 steps vs the oracle."""
    from oracle import oracle as O
    model = GameOfLife(Life512Experiment)
    start = model.gpu.arrays.cells.get()[:model.cells_num]
    field = O.Field((512, 512), 8, np.uint8, start)
    want = O.lifelike(field, O.moore_deltas(2), O.moore_rev(2), 8, 12,
                      NUM_STEPS, colors=False, hashes=True)
    for step in range(NUM_STEPS):
        model.step()
        assert cells_crc(model) == want[step], "step %d" % (step + 1)
