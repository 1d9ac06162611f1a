"""
Conway's Game of Life and relatives, with their experiments.

The models restate ``examples/game_of_life.py`` of the reference in the
xentica_amd DSL.  Their rules, seeds and experiment constants must be the
same for the reference's golden checksums to apply (for instance 1000
steps of :class:`GameOfLife` on :class:`GOLExperiment` hash to
2981695958, ``tests/core/test_base.py:41``).
"""
from xentica import core
from xentica import seeds
from xentica.core import color_effects
from xentica.tools.rules import LifeLike


class GameOfLife(core.CellularAutomaton):
    """B3/S23 on a 2-D torus with a single 1-bit ``state``."""

    state = core.IntegerProperty(max_val=1)

    class Topology:
        """2-D orthogonal lattice, Moore neighbourhood, torus."""

        dimensions = 2
        lattice = core.OrthogonalLattice()
        neighborhood = core.MooreNeighborhood()
        border = core.TorusBorder()

    def emit(self):
        """Every buffer carries the cell's state."""
        for buf in self.buffers:
            buf.state = self.main.state

    def absorb(self):
        """Count live neighbours; born on 3, survive on 2 or 3."""
        neighbors_alive = core.IntegerVariable()
        for i in range(len(self.buffers)):
            neighbors_alive += self.neighbors[i].buffer.state
        is_born = (8 >> neighbors_alive) & 1
        is_sustain = (12 >> neighbors_alive) & 1
        self.main.state = is_born | is_sustain & self.main.state

    @color_effects.MovingAverage
    def color(self):
        """Live cells white, dead cells black."""
        red = self.main.state * 255
        green = self.main.state * 255
        blue = self.main.state * 255
        return (red, green, blue)


class GameOfLifeStatic(GameOfLife):
    """Life surrounded by a static border of live cells."""

    class Topology(GameOfLife.Topology):
        """Same topology with a static border."""

        border = core.StaticBorder(1)


class GameOfLifeColor(GameOfLife):
    """Life whose live cells carry an RGB colour summed from neighbours."""

    state = core.IntegerProperty(max_val=1)
    red = core.IntegerProperty(max_val=255)
    green = core.IntegerProperty(max_val=255)
    blue = core.IntegerProperty(max_val=255)

    def emit(self):
        """Every buffer carries all four properties."""
        for buf in self.buffers:
            buf.state = self.main.state
            buf.red = self.main.red
            buf.green = self.main.green
            buf.blue = self.main.blue

    def absorb(self):
        """Life rule, then colour = sum over neighbours of (channel + 1)."""
        GameOfLife.absorb(self)
        red_sum = core.IntegerVariable()
        green_sum = core.IntegerVariable()
        blue_sum = core.IntegerVariable()
        for i in range(len(self.buffers)):
            red_sum += self.neighbors[i].buffer.red + 1
            green_sum += self.neighbors[i].buffer.green + 1
            blue_sum += self.neighbors[i].buffer.blue + 1
        self.main.red = red_sum * self.main.state
        self.main.green = green_sum * self.main.state
        self.main.blue = blue_sum * self.main.state

    @color_effects.MovingAverage
    def color(self):
        """The carried colour of live cells."""
        red = self.main.state * self.main.red
        green = self.main.state * self.main.green
        blue = self.main.state * self.main.blue
        return (red, green, blue)


class GameOfLife6D(GameOfLife):
    """Life in 6-D with a von Neumann neighbourhood."""

    class Topology(GameOfLife.Topology):
        """Six dimensions, von Neumann neighbourhood."""

        dimensions = 6
        neighborhood = core.VonNeumannNeighborhood()


class LifelikeCA(GameOfLife):
    """Life-like rule given as an interactive parameter."""

    rule = core.Parameter(
        default=LifeLike.golly2int("B3/S23"),
        interactive=True,
    )

    def absorb(self):
        """Birth bits 0-8, survival bits 9-17 of ``rule``."""
        neighbors_alive = core.IntegerVariable()
        for i in range(len(self.buffers)):
            neighbors_alive += self.neighbors[i].buffer.state
        is_born = (self.rule >> neighbors_alive) & 1
        is_sustain = (self.rule >> 9 >> neighbors_alive) & 1
        self.main.state = is_born | is_sustain & self.main.state

    def step(self):
        """Switch to B3/S23 at timestep 23."""
        if self.timestep == 23:
            self.rule = LifeLike.golly2int("B3/S23")
        super().step()


class GOLExperiment(core.Experiment):
    """A 100x100 random square in the middle of a 640x360 board."""

    word = "OBEY XENTICA"
    size = (640, 360, )
    zoom = 3
    pos = [0, 0]
    seed = seeds.patterns.BigBang(
        pos=(320, 180),
        size=(100, 100),
        vals={
            "state": seeds.random.RandInt(0, 1),
        }
    )


class GOLExperiment2(GOLExperiment):
    """The whole board random."""

    word = "XENTICA IS YOUR GODDESS"
    seed = seeds.patterns.PrimordialSoup(
        vals={
            "state": seeds.random.RandInt(0, 1),
        }
    )


class GOLExperimentColor(GOLExperiment):
    """Random states and colours; colours fade out slowly."""

    fade_in = 255
    fade_out = 10
    smooth_factor = 1
    seed = seeds.patterns.PrimordialSoup(
        vals={
            "state": seeds.random.RandInt(0, 1),
            "red": seeds.random.RandInt(0, 255),
            "green": seeds.random.RandInt(0, 255),
            "blue": seeds.random.RandInt(0, 255),
        }
    )


class GOLExperiment6D(GOLExperiment2):
    """Two spatial axes and four 3-cell micro-axes."""

    size = (640, 360, 3, 3, 3, 3)


class DiamoebaExperiment(GOLExperiment):
    """Starts with B35678/S5678 (``LifelikeCA`` switches rule at t=23)."""

    rule = LifeLike.golly2int("B35678/S5678")
