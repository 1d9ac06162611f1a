"""
Benchmark and parity workloads, sized as the project's measurement plan.

* :class:`Life512Experiment` -- Life B3/S23 on a 512x512 torus, 50 %
  random state from seed 42, used for per-step hash parity.
* :class:`Life16kExperiment` -- the same at 16384x16384 (the headline
  throughput workload, run by the bit-packed engine).
"""
from xentica_amd import core
from xentica_amd import seeds


class Life512Experiment(core.Experiment):
    """512x512, ``RandInt(0, 1)`` everywhere, word 42."""

    word = 42
    size = (512, 512)
    zoom = 1
    pos = [0, 0]
    seed = seeds.patterns.PrimordialSoup(
        vals={"state": seeds.random.RandInt(0, 1)})


class Life16kExperiment(Life512Experiment):
    """16384x16384, ``RandInt(0, 1)`` everywhere, word 42."""

    size = (16384, 16384)
