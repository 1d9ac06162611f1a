"""
The reference's ConservedLife (2-D and 4-D) and PoissonWalk examples on
the GPU.  The reference holds no golden hash for them (PoissonWalk's emit
even races on its neighbours' energy there); their invariant is exact:
the total energy never changes (the reference's own
``test_energy_conservation``, ``examples/conserved_life.py:237-256``).
"""
import numpy as np
import pytest

from examples.conserved_life import (
    ConservedLife, ConservedLife4D, BigBangExperiment, BigBangExperiment4D)
from examples.poisson_walk import PoissonWalk, GlyphExperiment
from xentica_amd.runtime.engines import PlainEngine

pytestmark = pytest.mark.gpu


def energy(model, bits):
    cells = model.gpu.arrays.cells.get()[:model.cells_num]
    return cells & np.uint64((1 << bits) - 1)


@pytest.mark.parametrize("cls,exp,bits,steps", [
    (ConservedLife, BigBangExperiment, 14, 300),
    (ConservedLife4D, BigBangExperiment4D, 14, 300),
    (PoissonWalk, GlyphExperiment, 16, 300)])
def test_total_energy_conserved(cls, exp, bits, steps):
    model = cls(exp)
    assert isinstance(model._engine, PlainEngine)
    start = energy(model, bits)
    total = int(start.sum(dtype=np.uint64))
    assert total > 0
    moved = False
    for block in range(steps // 50):
        for _ in range(50):
            model.step()
        now = energy(model, bits)
        assert int(now.sum(dtype=np.uint64)) == total, (block, model.timestep)
        moved = moved or not np.array_equal(now, start)
    assert moved, "the energy never moved"
    assert model.timestep == steps


def test_reference_benchmark_script_runs():
    """The reference's benchmark (examples/benchmark.py) on both engines:
    every model reports a rate; the default engine is the bit-sliced one
    for Conway's Life and the fused one for the other three models."""
    from examples import benchmark
    for engine in (None, "plain"):
        results = benchmark.run(num_steps=5, engine=engine, log=None)
        assert list(results) == [name for name, _, _ in benchmark.MODELS]
        for name, (speed, cells, used) in results.items():
            assert speed > 0 and cells == 640 * 360
            default = "bits" if name == "Conway's Life" else "fused"
            assert used == (engine or default), name
