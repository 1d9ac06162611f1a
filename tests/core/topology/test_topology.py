"""Topology (``tests/core/topology/test_*.py`` of the reference)."""
import pytest

from xentica_amd.core.exceptions import XenticaException
from xentica_amd.core.topology.border import TorusBorder
from xentica_amd.core.topology.lattice import OrthogonalLattice
from xentica_amd.core.topology.neighborhood import (
    MooreNeighborhood, VonNeumannNeighborhood)
from examples.game_of_life import GameOfLife, GOLExperiment


class GOLExperimentBroken(GOLExperiment):
    size = (23, )


@pytest.mark.parametrize("cls", [TorusBorder, OrthogonalLattice,
                                 MooreNeighborhood])
def test_incorrect_dimensions(cls):
    with pytest.raises(XenticaException):
        cls().dimensions = 0


def test_wrong_field_size():
    with pytest.raises(XenticaException):
        GameOfLife(GOLExperimentBroken)


def test_moore_offsets_and_pairing():
    """``neighborhood.py:159-167`` including the 3-D centre quirk."""
    hood = MooreNeighborhood()
    hood.dimensions = 2
    assert len(hood) == 8
    assert hood.neighbor_deltas[0] == (-1, -1)
    assert hood.rev_index_map == list(range(8))[::-1]
    hood3 = MooreNeighborhood()
    hood3.dimensions = 3
    assert len(hood3) == 26
    assert hood3.neighbor_deltas[13] == (0, 0, 0)
    assert (1, 1, 1) not in hood3.neighbor_deltas


def test_von_neumann():
    hood = VonNeumannNeighborhood()
    hood.dimensions = 3
    assert hood.neighbor_deltas == [
        (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    assert hood.rev_index_map == [1, 0, 3, 2, 5, 4]


def test_no_topology():
    from xentica_amd.core.base import CellularAutomaton
    with pytest.raises(XenticaException):
        class NoTopologyCA(CellularAutomaton):
            pass


def test_empty_topology():
    from xentica_amd.core.base import CellularAutomaton
    with pytest.raises(XenticaException):
        class NoLatticeCA(CellularAutomaton):
            class Topology:
                pass
