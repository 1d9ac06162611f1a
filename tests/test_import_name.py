"""
The reference's import name: ``xentica.<module>`` is the module
``xentica_amd.<module>`` (``xentica/core/__init__.py`` of the reference),
so user code written for Xentica runs unedited, and objects imported
under either name are the same objects.
"""
import binascii

import xentica_amd
from xentica import core, seeds
from xentica.core import CellularAutomaton, color_effects
from xentica.core.topology.neighborhood import MooreNeighborhood
from xentica.seeds.random import RandInt
from xentica.tools.rules import LifeLike
from xentica.utils.formatters import sizeof_fmt


def test_modules_are_shared():
    import xentica_amd.core.base
    import xentica.core.base
    assert core is xentica_amd.core
    assert xentica.core.base is xentica_amd.core.base
    assert CellularAutomaton is xentica_amd.core.base.CellularAutomaton
    assert MooreNeighborhood is core.MooreNeighborhood
    assert LifeLike.golly2int("B3/S23") == 6152
    assert sizeof_fmt(34567) == "34.57K"


class Life(core.CellularAutomaton):
    """The reference's Game of Life, written against ``xentica``."""

    state = core.IntegerProperty(max_val=1)

    class Topology:
        dimensions = 2
        lattice = core.OrthogonalLattice()
        neighborhood = core.MooreNeighborhood()
        border = core.TorusBorder()

    def emit(self):
        for i in range(len(self.buffers)):
            self.buffers[i].state = self.main.state

    def absorb(self):
        neighbors_alive = core.IntegerVariable()
        for i in range(len(self.buffers)):
            neighbors_alive += self.neighbors[i].buffer.state
        is_born = (8 >> neighbors_alive) & 1
        is_sustain = (12 >> neighbors_alive) & 1
        self.main.state = is_born | is_sustain & self.main.state

    @color_effects.MovingAverage
    def color(self):
        red = self.main.state * 255
        green = self.main.state * 255
        blue = self.main.state * 255
        return (red, green, blue)


class Experiment(core.Experiment):
    word = "OBEY XENTICA"
    size = (640, 360, )
    zoom = 3
    pos = [0, 0]
    seed = seeds.patterns.BigBang(
        pos=(320, 180), size=(100, 100), vals={"state": RandInt(0, 1)})


def test_a_model_written_against_xentica():
    """The golden CRC of ``tests/core/test_base.py:28-41`` from a model
    and experiment that import ``xentica`` only (host rendering of its
    generated code; the GPU engines run the same code)."""
    from xentica_amd.runtime.hostc import PlainHost
    model = Life(Experiment)
    assert model._engine.name == "bits"
    host = PlainHost(model)
    host.load(model.gpu.arrays.cells.get(), model.gpu.arrays.colors.get())
    host.step(1000)
    assert binascii.crc32(host.cells) == 2981695958
