"""
Sliding sand: a conserving sand-height rule whose coarse field moves half
a block in 16 steps and reshapes on the way.

Each cell holds a height 0-8 and a phase bit shared by the whole board (it
starts at 0 everywhere and flips every step).  On odd steps the heights
move one cell diagonally, (+1, +1): every cell takes its up-left
neighbour's height, as ``ShiftingSands`` does in the other direction
(``examples/shifting_sands.py`` of the reference).  On even steps sand
slides down steep edges along the same two axes: a cell at least three
grains higher than its right (lower) neighbour gives it one grain.  Every
grain given is taken by exactly one neighbour, so the total height is
conserved; a cell that gives has at least three grains and one that takes
at most five, so heights stay within 0-8.

Both the giving and the taking side of a slide are decided from heights
the cells exchange through their buffers (every buffer carries the cell's
state, as in Life), so ``emit`` is a plain copy and ``absorb`` alone
computes the new height.  The 16-step coarse field is neither the present
one nor the field rolled by one block: the drift is half a block, which
only the sub-block structure predicts, and the slides move mass further
where the field is steep.  It is the exact model the neural-CA surrogate
learns to predict.
"""
from xentica import core
from xentica import seeds
from xentica.core import color_effects

from examples.base import RegularCA


class SandSlide(RegularCA):
    """Half-speed diagonal drift with slides down steep edges."""

    grains = core.IntegerProperty(max_val=8)
    phase = core.IntegerProperty(max_val=1)

    def emit(self):
        """Every buffer carries the cell's height and phase."""
        for buf in self.buffers:
            buf.grains = self.main.grains
            buf.phase = self.main.phase

    def absorb(self):
        """Drift on odd steps, slides on even ones."""
        # Moore offsets: 0 = (-1, -1), 1 = (-1, 0), 3 = (0, -1),
        # 4 = (0, +1), 6 = (+1, 0)
        here = core.IntegerVariable()
        here += self.main.grains
        left = core.IntegerVariable()
        left += self.neighbors[1].buffer.grains
        right = core.IntegerVariable()
        right += self.neighbors[6].buffer.grains
        up = core.IntegerVariable()
        up += self.neighbors[3].buffer.grains
        down = core.IntegerVariable()
        down += self.neighbors[4].buffer.grains
        slid = core.IntegerVariable()
        slid += here
        slid -= here >= right + 3
        slid -= here >= down + 3
        slid += left >= here + 3
        slid += up >= here + 3
        drifted = core.IntegerVariable()
        drifted += self.neighbors[0].buffer.grains
        odd = core.IntegerVariable()
        odd += self.main.phase
        self.main.grains = odd * drifted + (1 - odd) * slid
        self.main.phase = 1 - odd

    @color_effects.ColorEffect
    def color(self):
        """Grey levels of the height."""
        val = self.main.grains * 30
        return (val, val, val)


class SandSlideExperiment(core.Experiment):
    """4096x4096 torus, heights uniform 0-8, phase 0, word 42."""

    word = 42
    size = (4096, 4096)
    zoom = 1
    pos = [0, 0]
    seed = seeds.patterns.PrimordialSoup(vals={
        "grains": seeds.random.RandInt(0, 8),
        "phase": 0,
    })


def slide_experiment(word, size=(4096, 4096)):
    """:class:`SandSlideExperiment` with another word and size."""
    return type("SandSlide%s" % word, (SandSlideExperiment, ),
                {"word": word, "size": tuple(size)})
