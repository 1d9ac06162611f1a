"""
Shifting sands: a non-uniform buffer interaction (``examples/shifting_sands.py``).

Each cell sends its whole value through ``buffers[0]`` only and takes the
sum of what it receives, so the pattern moves by (-1, -1) every step.
1000 steps on :class:`ShiftingSandsExperiment` hash to 1639048270
(``tests/core/test_base.py:182-189``).
"""
from xentica import core
from xentica import seeds
from xentica.core import color_effects

from examples.base import RegularCA, RegularExperiment


class ShiftingSands(RegularCA):
    """Emit the value in one direction, absorb by summing."""

    state = core.IntegerProperty(max_val=1)

    def emit(self):
        """The whole value goes to ``buffers[0]``; the rest carry 0."""
        direction = 0
        for i, buf in enumerate(self.buffers):
            if i == direction:
                buf.state = self.main.state
            else:
                buf.state = 0

    def absorb(self):
        """Sum everything received."""
        new_val = core.IntegerVariable()
        for i in range(len(self.buffers)):
            new_val += self.neighbors[i].buffer.state
        self.main.state = new_val

    @color_effects.MovingAverage
    def color(self):
        """Black and white."""
        red = self.main.state * 255
        green = self.main.state * 255
        blue = self.main.state * 255
        return (red, green, blue)


class ShiftingSandsExperiment(RegularExperiment):
    """A random square in the middle of the board."""

    word = "LET IT SHIFT"
    seed = seeds.patterns.BigBang(
        pos=(320, 180),
        size=(100, 100),
        vals={
            "state": seeds.random.RandInt(0, 1),
        }
    )
