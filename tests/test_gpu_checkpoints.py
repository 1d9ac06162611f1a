"""
Checkpoints of the packed fields and the batches in the reference's format
(``xentica/core/base.py:755-771``): a run interrupted by ``save()`` and
continued after ``load()`` -- into the same kind of object or across a
model and a packed field -- ends where the uninterrupted run does
(``tests/core/test_base.py:92-104``: golden CRC 2981695958 after 1000
steps of ``GameOfLife``/``GOLExperiment``).
"""
import binascii

import numpy as np
import pytest

from examples.game_of_life import GameOfLife, GOLExperiment
from examples.workloads import (
    LifeLikeRule, LifeLike3D, Life3DExperiment, batch_experiments)
from xentica_amd.runtime import kernels
from xentica_amd.runtime.bits_batch import BitsBatch
from xentica_amd.runtime.fused import Batch
from xentica_amd.runtime.lifebits import LifeBitsField, Life3DField, Life3DBytes

pytestmark = pytest.mark.gpu


def run(model, steps):
    for _ in range(steps):
        model.step()
    return model


def crc(model):
    return binascii.crc32(model.gpu.arrays.cells.get()[:model.cells_num])


def test_model_to_packed_field_and_back(tmp_path):
    first, second = str(tmp_path / "a.ca"), str(tmp_path / "b.ca")
    model = run(GameOfLife(GOLExperiment), 300)
    model.save(first)
    field = LifeBitsField(640, 360, 8, 12, depth=8)
    field.load(first)
    field.run(400)
    field.save(second)
    again = GameOfLife(GOLExperiment)
    again.load(second)
    assert crc(run(again, 300)) == 2981695958


def test_3d_fields_round_trip(tmp_path):
    path = str(tmp_path / "c.ca")

    class Small(Life3DExperiment):
        size = (64, 24, 16)
        engine = "plain"

    model = LifeLike3D(Small)
    rule = (Small.born, Small.sustain)
    cells = model.gpu.arrays.cells.get()[:model.cells_num]
    want = Life3DField(Small.size, *rule)
    want.set_cells(cells)
    want.step(10)
    for cls in (Life3DField, Life3DBytes):
        field = cls(Small.size, *rule)
        field.set_cells(cells)
        field.step(4)
        field.save(path)
        other = cls(Small.size, *rule)
        other.load(path)
        other.step(6)
        assert other.crc32() == want.crc32(), cls.__name__
    # a model loads the packed field's checkpoint
    run(model, 4)
    loaded = LifeLike3D(Small)
    loaded.load(path)
    assert np.array_equal(loaded.gpu.arrays.cells.get()[:model.cells_num],
                          model.gpu.arrays.cells.get()[:model.cells_num])


@pytest.mark.parametrize("kind", ["bits", "fused"])
def test_batch_round_trip(tmp_path, kind):
    path = str(tmp_path / ("batch_%s.ca" % kind))
    exps = batch_experiments(count=6, seed=13, side_min=40, side_max=300)

    def make(models):
        if kind == "bits":
            return BitsBatch(models, depth=6, strip_rows=16)
        return Batch(models, tile=(64, 16), depth=6)

    whole = make([LifeLikeRule(exp) for exp in exps])
    whole.run(50)
    whole.synchronize()
    whole.scatter()
    first = make([LifeLikeRule(exp) for exp in exps])
    first.run(20)
    first.save(path)
    second = make([LifeLikeRule(exp) for exp in exps])
    second.load(path)
    second.run(30)
    second.synchronize()
    second.scatter()
    for a, b in zip(whole.models, second.models):
        want = kernels.crc32_u8(a.gpu.arrays.cells.plane_ptr(a.cells_num),
                                a.cells_num)
        got = kernels.crc32_u8(b.gpu.arrays.cells.plane_ptr(b.cells_num),
                               b.cells_num)
        assert got == want
        assert np.array_equal(a.gpu.arrays.colors.get(),
                              b.gpu.arrays.colors.get())
