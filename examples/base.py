"""Shared templates of the example models (``examples/base.py`` of the reference)."""
import abc

from xentica import core


class RegularCA(core.CellularAutomaton):
    """2-D Moore torus template without logic."""

    class Topology:
        """2-D Moore neighbourhood on a torus."""

        dimensions = 2
        lattice = core.OrthogonalLattice()
        neighborhood = core.MooreNeighborhood()
        border = core.TorusBorder()

    @abc.abstractmethod
    def emit(self):
        """Emit phase logic."""

    @abc.abstractmethod
    def absorb(self):
        """Absorb phase logic."""

    @abc.abstractmethod
    def color(self):
        """Colouring logic."""


class RegularExperiment(core.Experiment):
    """640x360 board, zoom 3."""

    word = "I AM MUNDANE"
    size = (640, 360, )
    zoom = 3
    pos = [0, 0]
