"""
Noise TV: a per-cell LCG (``examples/noisetv.py``).

Absorb advances the cell's ``RandomProperty`` once and the colour reads
it again, so the generator advances twice per step.  100 steps on
:class:`NoiseTVExperiment` hash to 2773894957
(``tests/core/test_properties.py:64-71``).
"""
from xentica import core
from xentica import seeds
from xentica.core import color_effects
from xentica.seeds.random import RandInt
from xentica.tools import xmath

from examples.base import RegularCA, RegularExperiment


class NoiseTV(RegularCA):
    """Pseudorandom noise."""

    rng = core.RandomProperty()

    def emit(self):
        """Nothing to send: the noise is per cell."""

    def absorb(self):
        """Advance the generator."""
        return self.main.rng

    @color_effects.MovingAverage
    def color(self):
        """Grey level from the next generator value."""
        val = xmath.int(self.main.rng.uniform * 255)
        return (val, val, val)


class NoiseTVExperiment(RegularExperiment):
    """The whole board seeded with 16-bit values."""

    word = "PSHSHSHSH!"
    seed = seeds.patterns.PrimordialSoup(
        vals={
            "rng": RandInt(0, 2 ** 16 - 1)
        }
    )
