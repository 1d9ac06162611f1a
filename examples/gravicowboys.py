"""
GraviCowboys: integer masses in a float gravity field (``examples/gravicowboys.py``).

Each cell moves one unit of energy toward its strongest-gravity
neighbour and the field relaxes to the mean of the neighbours' field plus
mass.  Exercises ``FloatProperty`` bit-casting: 100 steps of
:class:`GraviCowboysExperiment` hash to 1068195512
(``tests/core/test_properties.py:97-104``).
"""
from xentica import core
from xentica import seeds
from xentica.core import color_effects
from xentica.tools import xmath

from examples.base import RegularCA, RegularExperiment


class GraviCowboys(RegularCA):
    """Masses following a diffusing gravity field."""

    energy = core.IntegerProperty(max_val=2 ** 12 - 1)
    new_energy = core.IntegerProperty(max_val=2 ** 12 - 1)
    gravity = core.FloatProperty()

    def emit(self):
        """Send one unit of energy toward the lowest field value."""
        self.main.new_energy = 1 * self.main.energy
        max_gravity = core.FloatVariable()
        max_gravity += xmath.min(*[nbr.main.gravity for nbr in self.neighbors])
        energy_passed = core.IntegerVariable()
        energy_passed *= 0
        denergy = core.IntegerVariable()
        for i, buf in enumerate(self.buffers):
            fit_dir = max_gravity == self.neighbors[i].main.gravity
            has_energy = (self.main.energy - energy_passed) > 0
            denergy *= 0
            denergy += fit_dir * has_energy * 1
            energy_passed += denergy
            buf.energy = denergy
            buf.gravity = self.main.gravity + self.main.energy
        self.main.new_energy -= energy_passed

    def absorb(self):
        """Collect incoming energy, average the field."""
        energy_passed = core.IntegerVariable()
        for i in range(len(self.buffers)):
            energy_passed += self.neighbors[i].buffer.energy
        self.main.energy = self.main.new_energy + energy_passed

        new_gravity = core.FloatVariable()
        for i in range(len(self.buffers)):
            new_gravity += self.neighbors[i].buffer.gravity
        self.main.gravity = new_gravity / (len(self.buffers) * 1.001)

    @color_effects.MovingAverage
    def color(self):
        """Mass as red, field as green and blue."""
        red = self.main.energy * 255
        green = xmath.int(self.main.gravity * 108)
        blue = xmath.int(self.main.gravity * 108)
        return (red, green, blue)


class GraviCowboysExperiment(RegularExperiment):
    """Three small random masses."""

    word = "LET IT GROOVE"
    seed = seeds.patterns.BigBang(
        pos=(320, 180),
        size=(10, 10),
        vals={
            "energy": seeds.random.RandInt(0, 1) * 1,
            "gravity": 0,
        }
    )
    seed += seeds.patterns.BigBang(
        pos=(20, 80),
        size=(10, 10),
        vals={
            "energy": seeds.random.RandInt(0, 1) * 1,
            "gravity": 0,
        }
    )
    seed += seeds.patterns.BigBang(
        pos=(400, 80),
        size=(10, 10),
        vals={
            "energy": seeds.random.RandInt(0, 1) * 1,
            "gravity": 0,
        }
    )
