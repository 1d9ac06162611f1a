"""
EnergyLife on the GPU engines against the oracle.

Per-step CRC-32 of the 128-bit main plane and the colours on a 96x80
torus (plain and fused engines); the 4096x4096 workload a few steps
against the oracle; and conserved total energy over a longer run.
"""
import binascii

import numpy as np
import pytest

from oracle import oracle
from examples.energy_life import (
    EnergyLife, EnergyLifeExperiment, EnergyLifeSmall)

pytestmark = pytest.mark.gpu


def experiment(base, engine):
    return type(base.__name__ + engine.title(), (base, ), {"engine": engine})


def model_cells(model):
    return model.gpu.arrays.cells.get()[:model.cells_num]


@pytest.mark.parametrize("engine", ["plain", "fused", "fused_unfolded",
                                    "fused_generic"])
def test_small_per_step(engine):
    model = EnergyLife(experiment(EnergyLifeSmall, engine))
    field = oracle.energy_life_field(EnergyLifeSmall.size)
    hashes = oracle.energy_life(field, 40, hashes=True)
    for step in range(40):
        model.step()
        assert binascii.crc32(model_cells(model)) == hashes[step], step
    assert np.array_equal(model.gpu.arrays.colors.get(), field.col)


def test_fused_runs_several_steps_per_read():
    model = EnergyLife(experiment(EnergyLifeSmall, "fused"))
    field = oracle.energy_life_field(EnergyLifeSmall.size)
    for _ in range(37):
        model.step()
    oracle.energy_life(field, 37)
    assert model_cells(model).tobytes() == field.cells.tobytes()
    assert np.array_equal(model.gpu.arrays.colors.get(), field.col)


def test_full_size_against_oracle():
    model = EnergyLife(EnergyLifeExperiment)
    field = oracle.energy_life_field(EnergyLifeExperiment.size)
    for _ in range(6):
        model.step()
    oracle.energy_life(field, 6, colors=False)
    assert model_cells(model).tobytes() == field.cells.tobytes()


def test_full_size_energy_conserved():
    model = EnergyLife(EnergyLifeExperiment)
    total = oracle.energy_fields(model_cells(model))[0].sum()
    for _ in range(4):
        for _ in range(50):
            model.step()
        energy, _, mode, _ = oracle.energy_fields(model_cells(model))
        assert energy.sum() == total
        assert energy.max() <= 255
        assert 0.05 < mode.mean() < 0.95


def test_batch_of_uneven_fields():
    """128-bit words in one persistent launch: every instance = its oracle."""
    from xentica_amd.runtime.fused import Batch
    sizes = [(96, 80), (64, 200), (130, 48)]
    exps = [type("Energy%d" % idx, (EnergyLifeSmall, ),
                 {"size": size, "word": 42 + idx})
            for idx, size in enumerate(sizes)]
    models = [EnergyLife(exp) for exp in exps]
    batch = Batch(models, tile=(16, 16), depth=2)
    batch.run(11)
    batch.synchronize()
    batch.scatter()
    for model, exp in zip(models, exps):
        field = oracle.energy_life_field(exp.size, exp.word)
        oracle.energy_life(field, 11)
        assert model_cells(model).tobytes() == field.cells.tobytes(), exp.size
        assert np.array_equal(model.gpu.arrays.colors.get(), field.col)
