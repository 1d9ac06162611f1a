"""
PoissonWalk: every cell a Poisson process passing energy to a neighbour.

The model of the reference's ``examples/poisson_walk.py`` (``emit`` /
``absorb`` / ``color`` at lines 35-126), restated with this package's DSL.
In the reference, ``emit`` rewrites a cell's energy while its neighbours
read that energy (their "valency", energy mod 8): the result depends on
thread timing there, and so it does on the plain engine here, which is the
one that runs it (emit stores the main state).  What does not depend on
timing is the total energy, since whatever a cell gives one neighbour
arrives there; the GPU tests check that.

* Emit: with probability ``1 - exp(-interval / (256 - min(255, energy)))``
  a cell with energy fires through its gate: it passes ``min(default
  energy, energy)`` (one less on a rare mutation) when the target's
  valency is non-zero or its own is zero, plus its own valency when the
  target's is zero; the gate advances by one each step, the interval
  counts the steps since the last transfer.
* Absorb: energy and gate offsets received are added.
"""
from xentica import core
from xentica import seeds
from xentica.core import color_effects
from xentica.tools import xmath


class PoissonWalk(core.CellularAutomaton):
    """Energy walkers on a 2-D Moore torus."""

    energy = core.IntegerProperty(max_val=2 ** 16 - 1)
    gate = core.IntegerProperty(max_val=2 ** 3 - 1)
    interval = core.IntegerProperty(max_val=2 ** 12 - 1)
    rng = core.RandomProperty()
    default_denergy = core.Parameter(default=1)

    class Topology:
        """2-D Moore neighbourhood on a torus."""

        dimensions = 2
        lattice = core.OrthogonalLattice()
        neighborhood = core.MooreNeighborhood()
        border = core.TorusBorder()

    def emit(self):
        """Fire (or not) through the gate."""
        main = self.main
        shade = xmath.min(255, main.energy)
        rate = xmath.float(main.interval) / xmath.float(256 - shade)
        chance = 1 - xmath.exp(-rate)
        fired = core.IntegerVariable()
        fired += xmath.int(main.rng.uniform < chance)
        fired *= main.energy > 0
        amount = xmath.min(self.default_denergy, shade)
        gate = main.gate
        mutated = core.IntegerVariable()
        mutated += xmath.int(main.rng.uniform < 0.0001)
        amount = amount - 1 * (amount > 0) * mutated
        passed = core.IntegerVariable()
        for i, buf in enumerate(self.buffers):
            theirs = self.neighbors[i].main.energy % 8
            mine = main.energy % 8
            through = (i == gate)
            full = amount * (theirs != 0 | mine == 0)
            potent = mine * (theirs == 0)
            passed += (full + potent) * fired * through
            buf.energy = passed * fired * through
            buf.gate = (main.gate + 7) * fired * through
        main.interval = (main.interval + 1) * (passed == 0)
        main.energy -= passed * (passed > 0)
        main.gate = (main.gate + 1) % 8

    def absorb(self):
        """Add what the neighbours sent."""
        energy_in = core.IntegerVariable()
        gate_in = core.IntegerVariable()
        for i in range(len(self.buffers)):
            energy_in += self.neighbors[i].buffer.energy
            gate_in += self.neighbors[i].buffer.gate
        self.main.energy += energy_in
        self.main.gate += gate_in % 8

    @color_effects.MovingAverage
    def color(self):
        """Valency (energy mod 8) as RGB bits; valency 0 as grey."""
        energy = xmath.min(255, self.main.energy)
        vacuum = (energy % 8 == 0) * energy
        red = (((energy % 8) >> 2) & 1) * 255 + vacuum
        blue = (((energy % 8) >> 1) & 1) * 255 + vacuum
        green = ((energy % 8) & 1) * 255 + vacuum
        return (red, green, blue)


class GlyphExperiment(core.Experiment):
    """640x360: 18 squares of random energy on a field of random gates."""

    word = "PSHSHSHSH!"
    size = (640, 360, )
    zoom = 3
    pos = [0, 0]
    fade_in = 1
    fade_out = 1
    smooth_factor = 100
    speed = 200
    seed = seeds.patterns.PrimordialSoup(vals={
        "energy": 0,
        "gate": seeds.random.RandInt(0, 2 ** 3 - 1),
        "interval": seeds.random.RandInt(0, 2 ** 12 - 1),
        "rng": seeds.random.RandInt(0, 2 ** 16 - 1)})
    for _i in range(18):
        seed += seeds.patterns.BigBang(
            pos=(30 + (_i % 6) * 100, 40 + (_i // 6) * 100), size=(80, 80),
            vals={"energy": seeds.random.RandInt(0, 2 ** 7 - 1),
                  "gate": seeds.random.RandInt(0, 2 ** 3 - 1),
                  "interval": seeds.random.RandInt(0, 2 ** 12 - 1),
                  "rng": seeds.random.RandInt(0, 2 ** 16 - 1)})
