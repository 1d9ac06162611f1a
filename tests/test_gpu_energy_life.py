"""
EnergyLife on the GPU engines against the oracle.

Per-step CRC-32 of the main plane (64-bit words, and the 128-bit variant)
and the colours on a 96x80 torus (plain and fused engines); the 4096x4096
workload a few steps against the oracle; conserved total energy over a
longer run, and after every one of the workload's 5000 steps (summed on
the device).
"""
import binascii

import numpy as np
import pytest

from oracle import oracle
from examples.energy_life import (
    EnergyLife, EnergyLifeExperiment, EnergyLifeSmall, EnergyLifeWide,
    field_position, field_values)
from xentica_amd.runtime import kernels

pytestmark = pytest.mark.gpu


def experiment(base, engine):
    return type(base.__name__ + engine.title(), (base, ), {"engine": engine})


def model_cells(model):
    return model.gpu.arrays.cells.get()[:model.cells_num]


@pytest.mark.parametrize("wide", [False, True], ids=["64bit", "128bit"])
@pytest.mark.parametrize("engine", ["plain", "fused", "fused_unfolded",
                                    "fused_generic"])
def test_small_per_step(engine, wide):
    model = (EnergyLifeWide if wide else EnergyLife)(
        experiment(EnergyLifeSmall, engine))
    field = oracle.energy_life_field(EnergyLifeSmall.size, wide=wide)
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


def device_energy(model):
    """Total energy of the model's current main plane, summed on the
    device."""
    shift, width = field_position(model, "energy")
    model._engine.flush()
    cells = model.gpu.arrays.cells
    return kernels.field_sum(cells.plane_ptr(model.cells_num),
                             model.cells_num, cells.dtype.itemsize, shift,
                             width)


def test_device_field_sum_equals_host():
    """The device sum of each field equals NumPy's over the downloaded
    plane, before and after some steps."""
    model = EnergyLife(EnergyLifeExperiment)
    for steps in (0, 7):
        for _ in range(steps):
            model.step()
        model._engine.flush()
        cells = model.gpu.arrays.cells
        for name in ("energy", "genome", "mode", "rng"):
            shift, width = field_position(model, name)
            want = int(field_values(model, name).sum())
            got = kernels.field_sum(cells.plane_ptr(model.cells_num),
                                    model.cells_num, cells.dtype.itemsize,
                                    shift, width)
            assert got == want, (steps, name)


def test_energy_conserved_every_step_of_the_workload():
    """The C3 validation run: 5000 steps of EnergyLife 4096x4096 (one
    step per launch), the total energy after every step equal to the
    seed's."""
    model = EnergyLife(EnergyLifeExperiment)
    total = device_energy(model)
    assert total == int(field_values(model, "energy").sum())
    print("EnergyLife 4096^2, 5000 steps, energy %d" % total, flush=True)
    for step in range(5000):
        model.step()
        assert device_energy(model) == total, "after step %d" % (step + 1)


@pytest.mark.parametrize("wide", [False, True], ids=["64bit", "128bit"])
def test_batch_of_uneven_fields(wide):
    """64-bit (and 128-bit) words in one persistent launch: every instance
    equals its oracle."""
    from xentica_amd.runtime.fused import Batch
    sizes = [(96, 80), (64, 200), (130, 48)]
    exps = [type("Energy%d" % idx, (EnergyLifeSmall, ),
                 {"size": size, "word": 42 + idx})
            for idx, size in enumerate(sizes)]
    models = [(EnergyLifeWide if wide else EnergyLife)(exp) for exp in exps]
    batch = Batch(models, tile=(16, 16), depth=2)
    batch.run(11)
    batch.synchronize()
    batch.scatter()
    for model, exp in zip(models, exps):
        field = oracle.energy_life_field(exp.size, exp.word, wide=wide)
        oracle.energy_life(field, 11)
        assert model_cells(model).tobytes() == field.cells.tobytes(), exp.size
        assert np.array_equal(model.gpu.arrays.colors.get(), field.col)
