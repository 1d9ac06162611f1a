"""
Lattices (``xentica/core/topology/lattice.py``).

:class:`OrthogonalLattice` lays cells out with ``x0`` fastest: the flat
index is ``x0 + w0*x1 + w0*w1*x2 + ...`` (``lattice.py:189-204``).  Field
sizes reach the generated code as the constants ``_w0, _w1, ...``
(``lattice.py:48-58``); the plain kernels render them as ``#define``s,
the batched kernels as per-instance runtime values.
"""
import abc

from xentica_amd.core.mixins import DimensionsMixin, BscaDetectorMixin
from xentica_amd.core.variables import Constant
from xentica_amd.core.exceptions import XenticaException

__all__ = ['Lattice', 'OrthogonalLattice', ]


class Lattice(DimensionsMixin, BscaDetectorMixin, metaclass=abc.ABCMeta):
    """The base class for all lattices."""

    #: The prefix used in C code for field size constants.
    width_prefix = "_w"

    def _define_constants_once(self):
        """Register the field size constants with the model."""
        bsca = self.bsca
        num_dimensions = bsca.topology.dimensions
        size = getattr(bsca, "size", None)
        for i in range(num_dimensions):
            if size is None or i >= len(size):
                msg = "Wrong field's dimensionality ({} instead of {})."
                raise XenticaException(msg.format(
                    len(size) if size is not None else 0, num_dimensions))
            bsca.define_constant(
                Constant("%s%d" % (self.width_prefix, i), size[i]))

    @abc.abstractmethod
    def index_to_coord_code(self, index_name, coord_prefix):
        """Generate C code computing coordinates from a cell index."""

    @abc.abstractmethod
    def index_to_coord(self, idx, bsca):
        """Compute coordinates from a cell index (or index array) in Python."""

    @abc.abstractmethod
    def coord_to_index_code(self, coord_prefix):
        """Generate a C expression computing a cell index from coordinates."""

    @abc.abstractmethod
    def is_off_board_code(self, coord_prefix):
        """Generate a C boolean expression testing for off-board coords."""


class OrthogonalLattice(Lattice):
    """N-dimensional orthogonal lattice."""

    supported_dimensions = list(range(1, 100))

    def index_to_coord_code(self, index_name, coord_prefix):
        """
        Emit ``int _i = i;`` then peel coordinates off by ``%`` and ``/=``.

        Same statements as ``lattice.py:146-170``.
        """
        self._define_constants_once()
        width = self.width_prefix
        local = "_" + index_name
        lines = ["int %s = %s;\n" % (local, index_name)]
        last = self.dimensions - 1
        for axis in range(self.dimensions):
            if axis < last:
                lines.append("int %s%d = %s %% %s%d;\n" % (
                    coord_prefix, axis, local, width, axis))
                lines.append("%s /= %s%d;\n" % (local, width, axis))
            else:
                lines.append("int %s%d = %s;\n" % (coord_prefix, axis, local))
        return "".join(lines)

    def index_to_coord(self, idx, bsca):
        """Peel coordinates off ``idx`` in Python (``lattice.py:172-187``)."""
        coord = []
        last = self.dimensions - 1
        for axis in range(bsca.topology.dimensions):
            if axis < last:
                coord.append(idx % bsca.size[axis])
                idx //= bsca.size[axis]
            else:
                coord.append(idx)
        return coord

    def coord_to_index_code(self, coord_prefix):
        """Render ``x0 + _w0 * x1 + _w1 * _w0 * x2 ...`` (``lattice.py:189-204``)."""
        self._define_constants_once()
        summands = []
        for axis in range(self.dimensions):
            summand = "%s%d" % (coord_prefix, axis)
            for inner in range(axis):
                summand = "%s%d * %s" % (self.width_prefix, inner, summand)
            summands.append(summand)
        return " + ".join(summands)

    def is_off_board_code(self, coord_prefix):
        """Render the off-board test of ``lattice.py:206-221``."""
        self._define_constants_once()
        return " || ".join(
            "{x}{i} < 0 || {x}{i} >= {w}{i}".format(
                x=coord_prefix, i=axis, w=self.width_prefix)
            for axis in range(self.dimensions))
