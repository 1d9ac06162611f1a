"""
Seed patterns (``xentica/seeds/patterns.py``).

A pattern fills the step-0 cells array through ``generate(cells, bsca)``.
Values are drawn key by key in ``sorted()`` order and packed with the
model's ``pack_state`` (no masking, ``base.py:510-524``).

* :class:`BigBang` fills an N-cube area; missing size/position are drawn
  from the stdlib stream and the area is clamped onto the board
  (``patterns.py:187-235``).  Values are assigned in flat index order.
* :class:`PrimordialSoup` fills the whole board (``patterns.py:238-274``).
* ``a + b`` chains patterns, applied left to right on one shared stream
  (``patterns.py:137-159``).
"""
import abc

import numpy as np

from .random import LocalRandom

__all__ = ['RandomPattern', 'BigBang', 'PrimordialSoup', 'ValDict', ]


class ValDictMeta(type):
    """A placeholder metaclass for :class:`ValDict`."""


class ValDict(metaclass=ValDictMeta):
    """
    A read-only dict whose descriptor values are resolved on access.

    :param dct: mixed values, possibly descriptors such as ``RandInt``
    :param parent: the object descriptors are resolved against
    """

    def __init__(self, dct, parent=None):
        """Initialize the class."""
        self._d = dct
        self.parent = parent
        if parent is None:
            self.parent = self
            self.random = LocalRandom()

    def items(self):
        """Iterate over resolved items."""
        for key in self._d.keys():
            yield key, self[key]

    def keys(self):
        """Iterate over keys."""
        yield from self._d.keys()

    def __getitem__(self, key):
        """Resolve descriptors through ``__get__``, return others as is."""
        if key in self._d:
            val = self._d[key]
            if hasattr(val, '__get__'):
                return val.__get__(self.parent, self.parent.__class__)
            return val
        raise KeyError(key)

    def __setitem__(self, key, val):
        """Direct item setting is not supported."""
        raise NotImplementedError


class RandomPattern:
    """The base class for random patterns."""

    def __init__(self, vals):
        """Initialize the class."""
        self._random = LocalRandom()
        self.vals = ValDict(vals, self)

    @property
    def random(self):
        """Get the random stream."""
        return self._random

    @random.setter
    def random(self, val):
        """Set the random stream."""
        self._random = val

    def __add__(self, other):
        """Chain two patterns."""
        return ChainedPattern(self, other)

    def _draw_state(self):
        """Draw every value, keys in sorted order."""
        return {name: self.vals[name] for name in sorted(self.vals.keys())}

    @abc.abstractmethod
    def generate(self, cells, bsca):
        """Write the pattern into ``cells`` (the flat state array)."""


class ChainedPattern(RandomPattern):
    """Two patterns applied one after the other."""

    def __init__(self, pattern1, pattern2):
        """Store the two patterns."""
        self._pattern1 = pattern1
        self._pattern2 = pattern2
        super().__init__({})

    @RandomPattern.random.setter
    def random(self, val):
        """Share the stream with both patterns."""
        self._pattern1.random = val
        self._pattern2.random = val

    def generate(self, cells, bsca):
        """Generate both patterns in order."""
        self._pattern1.generate(cells, bsca)
        self._pattern2.generate(cells, bsca)


class BigBang(RandomPattern):
    """
    A random N-cube area on an otherwise unchanged board.

    :param vals: mixed values, possibly descriptors
    :param pos: lowest corner of the area
    :param size: size of the area per axis
    """

    def __init__(self, vals, pos=None, size=None):
        """Initialize class."""
        self.pos = np.asarray(pos) if pos else None
        self.size = np.asarray(size) if size else None
        super().__init__(vals)

    def _prepare_area(self, bsca_size):
        """Draw missing size/position, then clamp the area onto the board."""
        dims = range(len(bsca_size))
        randint = self.random.standard.randint
        if self.size is None:
            self.size = np.asarray([randint(1, bsca_size[i]) for i in dims])
        if self.pos is None:
            self.pos = np.asarray([randint(0, bsca_size[i]) for i in dims])
        for i, coord in enumerate(self.pos):
            width, side = self.size[i], bsca_size[i]
            if coord + width >= side:
                self.pos[i] = side - width
            self.pos[i] = max(0, self.pos[i])

    def area_mask(self, bsca):
        """Boolean mask over the board's flat indices that lie in the area."""
        coords = bsca.index_to_coord(np.arange(0, bsca.cells_num))
        area = None
        for i in range(len(bsca.size)):
            inside = (coords[i] >= self.pos[i])
            inside &= (coords[i] < self.pos[i] + self.size[i])
            area = inside if area is None else area & inside
        return area

    def generate(self, cells, bsca):
        """Fill the area in flat index order."""
        self._prepare_area(bsca.size)
        area = self.area_mask(bsca)
        cells[np.where(area)] = bsca.pack_state(self._draw_state())


def _per_cell_draws(value):
    """Whether ``value`` is a number or an expression whose random leaves
    all draw one value per cell from the NumPy stream (no ``constant``
    draws from the stdlib stream)."""
    from .random import RandInt
    from xentica_amd.core.expressions import PatternExpression
    if isinstance(value, RandInt):
        return not value._constant  # pylint: disable=protected-access
    if isinstance(value, PatternExpression):
        return all(_per_cell_draws(arg) for arg in (value.arg1, value.arg2)
                   if arg is not None)
    return isinstance(value, (int, float, np.integer, np.floating))


class PrimordialSoup(RandomPattern):
    """
    Random values over the whole board.

    A board of more than ``CHUNK`` cells with a single field whose values
    are all drawn per cell is filled ``CHUNK`` cells at a time: the legacy
    ``RandomState`` stream continues across calls, so the values are the
    ones one draw of the whole board gives, without its int64 temporaries
    (8 bytes per cell for every intermediate value).
    """

    #: cells drawn at a time on large boards
    CHUNK = 1 << 24

    def __init__(self, vals):
        """Initialize class."""
        self.size = None
        super().__init__(vals)

    def generate(self, cells, bsca):
        """Fill every cell."""
        count = bsca.cells_num
        keys = list(self.vals.keys())
        if count > self.CHUNK and len(keys) == 1 and \
                _per_cell_draws(self.vals._d[keys[0]]):  # pylint: disable=protected-access
            for start in range(0, count, self.CHUNK):
                part = min(self.CHUNK, count - start)
                self.size = (part, )
                cells[start:start + part] = bsca.pack_state(
                    self._draw_state())
            self.size = bsca.size
            return
        self.size = bsca.size
        cells[:count] = bsca.pack_state(self._draw_state())
