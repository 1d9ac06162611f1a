"""
The reference's benchmark (``tests/benchmark.py``): four example models,
``NUM_STEPS`` calls of ``model.step()`` each from Python, and
``cells/s = NUM_STEPS * cells_num / seconds`` printed with ``sizeof_fmt``
(``tests/benchmark.py:19-36``).

Two differences, both so that the figure means what it says: the kernels
are compiled and launched once before the clock starts, and the clock
stops after the device has finished every step (the reference reads it
while its last launches may still be queued).  ``run()`` also returns the
figures, which ``bench.py --only refbench`` reports.

Usage: ``python -m examples.benchmark [steps] [engine]``.
"""
import sys
import time

from examples.game_of_life import GameOfLife, GOLExperiment
from examples.shifting_sands import ShiftingSands, ShiftingSandsExperiment
from examples.noisetv import NoiseTV, NoiseTVExperiment
from examples.evolife import EvoLife, BigBangExperiment
from xentica.utils.formatters import sizeof_fmt

MODELS = [
    ("Conway's Life", GameOfLife, GOLExperiment),
    ("Shifting Sands", ShiftingSands, ShiftingSandsExperiment),
    ("Noise TV", NoiseTV, NoiseTVExperiment),
    ("EvoLife", EvoLife, BigBangExperiment),
]
NUM_STEPS = 100000


def run(num_steps=NUM_STEPS, engine=None, log=print):
    """``{name: (cells/s, cells, engine)}`` of every model."""
    results = {}
    for name, model, experiment in MODELS:
        if engine:
            experiment = type(experiment.__name__ + engine.title(),
                              (experiment, ), {"engine": engine})
        ca = model(experiment)
        ca.step()
        ca.gpu.arrays.cells.get()
        start = time.perf_counter()
        for _ in range(num_steps):
            ca.step()
        ca.gpu.arrays.cells.get()   # deferred steps done, device finished
        seconds = time.perf_counter() - start
        speed = num_steps * ca.cells_num / seconds
        results[name] = (speed, ca.cells_num, ca._engine.name)
        if log:
            log("%s: %s cells/s" % (name, sizeof_fmt(speed)))
        del ca
    return results


if __name__ == "__main__":
    run(int(sys.argv[1]) if len(sys.argv) > 1 else NUM_STEPS,
        sys.argv[2] if len(sys.argv) > 2 else None)
