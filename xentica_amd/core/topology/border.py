"""
Borders (``xentica/core/topology/border.py``).

* :class:`TorusBorder`: when any coordinate of a neighbour is off the
  board, every coordinate becomes ``(c + w) % w`` (``border.py:104-117``,
  applied at ``properties.py:457-466``).
* :class:`StaticBorder`: the whole packed word of an off-board neighbour,
  main state and buffers alike, is the constant (``border.py:144-151``,
  ``properties.py:467-483``).
"""
import abc

from xentica_amd.core.mixins import DimensionsMixin

__all__ = [
    'Border', 'WrappedBorder', 'GeneratedBorder',
    'TorusBorder', 'StaticBorder',
]


class Border(DimensionsMixin):
    """The base class for all types of borders."""

    def __init__(self):
        """Initialize common things for all borders."""
        self.topology = None
        super().__init__()


class WrappedBorder(Border):
    """Borders mapping off-board coordinates back onto the board."""

    @abc.abstractmethod
    def wrap_coords(self, coord_prefix):
        """Generate C code translating off-board coordinates to on-board."""


class GeneratedBorder(Border):
    """Borders generating the state of off-board cells."""

    @abc.abstractmethod
    def off_board_state(self, coord_prefix):
        """Generate a C expression for the state of an off-board cell."""


class TorusBorder(WrappedBorder):
    """Wraps the field into an N-torus."""

    supported_dimensions = list(range(1, 100))

    def wrap_coords(self, coord_prefix):
        """Emit ``x<i> = (x<i> + _w<i>) % _w<i>;`` for every axis."""
        width = self.topology.lattice.width_prefix
        return "".join(
            "{x}{i} = ({x}{i} + {w}{i}) % {w}{i};\n".format(
                x=coord_prefix, i=axis, w=width)
            for axis in range(self.dimensions))


class StaticBorder(GeneratedBorder):
    """Surrounds the field with cells of one constant packed state."""

    supported_dimensions = list(range(1, 100))

    def __init__(self, value=0):
        """Store the static value."""
        self._value = value
        super().__init__()

    @property
    def value(self):
        """The packed word every off-board cell holds."""
        return self._value

    def off_board_state(self, coord_prefix):
        """Return the constant as C text."""
        return str(self._value)
