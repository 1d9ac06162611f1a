"""
Sand heights: ShiftingSands with integer heights 0-8 (4-bit state).

The rule is the reference's ShiftingSands (``examples/shifting_sands.py``
of the reference): every cell sends its value through ``buffers[0]`` only
and takes the sum of what it receives, so the field moves by (-1, -1)
every step and 16 steps move it by exactly one 16x16 block.  It is the
exact model the neural-CA surrogate learns to predict.
"""
from xentica import core
from xentica import seeds

from examples.shifting_sands import ShiftingSands


class SandHeights(ShiftingSands):
    """ShiftingSands on a 4-bit height."""

    state = core.IntegerProperty(max_val=8)


class SandHeightsExperiment(core.Experiment):
    """4096x4096 torus, heights uniform 0-8, word 42."""

    word = 42
    size = (4096, 4096)
    zoom = 1
    pos = [0, 0]
    seed = seeds.patterns.PrimordialSoup(vals={
        "state": seeds.random.RandInt(0, 8)})


def sand_experiment(word, size=(4096, 4096)):
    """:class:`SandHeightsExperiment` with another word and size."""
    return type("SandHeights%s" % word, (SandHeightsExperiment, ),
                {"word": word, "size": tuple(size)})
