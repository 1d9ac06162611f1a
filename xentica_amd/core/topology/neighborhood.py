"""
Neighbourhoods (``xentica/core/topology/neighborhood.py``).

Neighbour ``j`` of a cell sits at offset ``deltas[j]``.  The pairing of
buffer slots is ``rev_index_map``: a cell stores ``buffers[q]`` in slot
``rev[q] + 1`` and reads ``neighbors[j].buffer`` from slot ``j + 1`` of
the neighbour (``xentica/core/base.py:223-236``).

Moore offsets are ``itertools.product([-1, 0, 1], repeat=d)`` filtered by
``!= (0, 0)`` and cut to ``3**d - 1`` entries, exactly as
``neighborhood.py:159-167`` does.  In 2-D that is the eight true
neighbours; in 1-D and 3-D and higher the filter never matches, so the
centre cell stays in the list and the last offset ``(1, ..., 1)`` is the
one that is dropped.  That is the reference's behaviour and it is kept
here: every hash of a 3-D Moore model depends on it.
"""
import abc
import itertools

from xentica_amd.core.mixins import DimensionsMixin

__all__ = [
    'Neighborhood', 'OrthogonalNeighborhood',
    'MooreNeighborhood', 'VonNeumannNeighborhood',
]


class Neighborhood(DimensionsMixin):
    """The base class for all neighbourhood topologies."""

    def __init__(self):
        """Initialize main attributes."""
        self.num_neighbors = None
        self.topology = None
        self._delta2str = {-1: " - 1", 0: "", 1: " + 1"}
        self._neighbor_deltas = []
        self.rev_index_map = []
        super().__init__()

    def __len__(self):
        """Return the number of neighbours of a cell."""
        return self.num_neighbors or 0

    @property
    def neighbor_deltas(self):
        """Offsets of the neighbours, one tuple per neighbour index."""
        return list(self._neighbor_deltas[:len(self)])

    @abc.abstractmethod
    def neighbor_coords(self, index, coord_prefix, neighbor_prefix):
        """Generate C code computing a neighbour's coordinates."""

    @abc.abstractmethod
    def neighbor_state(self, neighbor_index, state_index, coord_prefix):
        """Generate a C expression reading a neighbour's state plane."""


class OrthogonalNeighborhood(Neighborhood):
    """Neighbourhoods on an orthogonal lattice (``neighborhood.py:107-148``)."""

    supported_dimensions = list(range(1, 100))

    def neighbor_coords(self, index, coord_prefix, neighbor_prefix):
        """Emit ``_nx<i> = _x<i> +/- 1;`` per axis."""
        delta = self._neighbor_deltas[index]
        return "".join(
            "%s%d = %s%d%s;\n" % (neighbor_prefix, axis, coord_prefix, axis,
                                  self._delta2str[delta[axis]])
            for axis in range(self.dimensions))

    def neighbor_state(self, neighbor_index, state_index, coord_prefix):
        """Render ``fld[<index of coords> + n * <plane>]``."""
        cell_index = self.topology.lattice.coord_to_index_code(coord_prefix)
        return "fld[%s + n * %s]" % (cell_index, state_index)


class MooreNeighborhood(OrthogonalNeighborhood):
    """N-dimensional Moore neighbourhood, with the reference's offset list."""

    @OrthogonalNeighborhood.dimensions.setter
    def dimensions(self, num_dim):
        """Set the number of neighbours and their offsets."""
        OrthogonalNeighborhood.dimensions.fset(self, num_dim)
        self.num_neighbors = 3 ** num_dim - 1
        self._neighbor_deltas = [
            delta for delta in itertools.product([-1, 0, 1], repeat=num_dim)
            if delta != (0, 0)
        ]
        self.rev_index_map = list(range(self.num_neighbors))[::-1]


class VonNeumannNeighborhood(OrthogonalNeighborhood):
    """N-dimensional von Neumann neighbourhood: ``+e0, -e0, +e1, -e1, ...``."""

    @OrthogonalNeighborhood.dimensions.setter
    def dimensions(self, num_dim):
        """Set the number of neighbours and their offsets."""
        OrthogonalNeighborhood.dimensions.fset(self, num_dim)
        self.num_neighbors = 2 * num_dim
        self._neighbor_deltas = []
        self.rev_index_map = []
        for axis in range(num_dim):
            for sign in (1, -1):
                self._neighbor_deltas.append(tuple(
                    sign if axis == other else 0 for other in range(num_dim)))
            self.rev_index_map += [axis * 2 + 1, axis * 2]
