"""
Cell state schema: properties and their bit packing.

Behaviour follows ``xentica/core/properties.py``:

* every property gets the smallest of 8/16/32/64-bit unsigned C types that
  holds its bit width, and its C variable has that type, so a store
  truncates before the pack masks it (``properties.py:70-151``);
* all top-level properties of a model are packed into ONE container word,
  first property in the lowest bits (``properties.py:496-541``);
* reading any field of a container loads and unpacks the whole word
  through a 32-bit ``int`` temporary (``properties.py:405-413, 496-514``);
* writing any field marks the container for a packed store at the end of
  the kernel body (``properties.py:415-429, 516-541``);
* :class:`RandomProperty` advances the 16-bit LCG
  ``r = (r * 58321 + 11113) % 65535`` on every access
  (``properties.py:622-660``);
* :class:`TotalisticRuleProperty` holds birth/survival bit masks
  (``properties.py:544-619``).

Beyond the reference: a container wider than 64 bits is one
``unsigned __int128`` word (host dtype :data:`UINT128`: two little-endian
uint64 halves, low half first), so a model whose fields sum to 65-128
bits keeps the one-word semantics above; and :class:`XorShiftProperty`
is a 32-bit xorshift generator with :class:`RandomProperty`'s
advance-on-access behaviour.

Where a container's word lives in memory is asked of the model's
addressing scheme (``bsca.addressing``): the plain scheme renders the
reference's ``fld[i + n * plane]`` and neighbour-coordinate code
verbatim, the tiled scheme of the fused kernels renders LDS accesses.
"""
import math
import abc
from collections import OrderedDict
from functools import cached_property

import numpy as np

from xentica_amd.core.expressions import DeferredExpression
from xentica_amd.core.exceptions import XenticaException
from xentica_amd.core.mixins import BscaDetectorMixin
from xentica_amd.tools import xmath

__all__ = [
    'Property', 'IntegerProperty', 'FloatProperty', 'ContainerProperty',
    'TotalisticRuleProperty', 'RandomProperty', 'XorShiftProperty',
    'UINT128',
]

#: Host dtype of a 128-bit container word: the little-endian byte image
#: of ``unsigned __int128`` (low 64 bits first).
UINT128 = np.dtype([('lo', '<u8'), ('hi', '<u8')])

#: (bit width, NumPy dtype, C type stem), smallest first.
_TYPES = (
    (8, np.uint8, 'char'),
    (16, np.uint16, 'short'),
    (32, np.uint32, 'int'),
    (64, np.uint64, 'long'),
)
#: The word of a container of 65-128 bits (containers only: a single
#: property wider than 64 bits keeps the reference's 64-bit type and a
#: ``width`` of several words, ``properties.py:82-99``).
_WIDE_TYPE = (128, UINT128, '__int128')


class Property(DeferredExpression, BscaDetectorMixin):
    """
    Base class for all properties.

    A property renders as its C variable name in expressions.  Nested
    properties stored on an instance behave as descriptors: reading one
    calls its ``__get__`` and assigning one calls its ``__set__``.
    """

    def __init__(self):
        """Initialize default attributes."""
        self._buf_num, self._nbr_num = 1, 1
        super().__init__()

    @cached_property
    def best_type(self):
        """``(bit_width, numpy_dtype, c_type_stem)`` that holds the property."""
        for entry in _TYPES:
            if self.bit_width <= entry[0]:
                return entry
        return _TYPES[-1]

    @cached_property
    def dtype(self):
        """NumPy dtype that stores the property."""
        return self.best_type[1]

    @cached_property
    def ctype(self):
        """C type that stores the property."""
        return 'unsigned ' + self.best_type[2]

    @property
    def _num_neighbors(self):
        """Number of neighbours of a cell in the current model."""
        return len(self.bsca.buffers)

    @cached_property
    def bit_width(self):
        """Number of bits the property occupies in the packed word."""
        return self.calc_bit_width()

    @cached_property
    def width(self):
        """Number of words of ``best_type`` the bit width spans."""
        return int(math.ceil(self.bit_width / self.best_type[0]))

    @abc.abstractmethod
    def calc_bit_width(self):
        """Calculate the property's bit width (default 1)."""
        return 1

    @property
    def buf_num(self):
        """Memory plane of the property (0 = main state)."""
        return self._buf_num

    @buf_num.setter
    def buf_num(self, val):
        self._buf_num = val

    @property
    def nbr_num(self):
        """Neighbour index of the property (-1 = the cell itself)."""
        return self._nbr_num

    @nbr_num.setter
    def nbr_num(self, val):
        self._nbr_num = val

    def __getattribute__(self, attr):
        """Resolve nested properties as descriptors."""
        obj = object.__getattribute__(self, attr)
        if isinstance(obj, Property) and attr != "__class__":
            return obj.__get__(self, type(self))
        return obj

    def __setattr__(self, attr, val):
        """Assign nested properties through their ``__set__``."""
        try:
            obj = object.__getattribute__(self, attr)
        except AttributeError:
            object.__setattr__(self, attr, val)
        else:
            if hasattr(obj, '__set__'):
                obj.__set__(self, val)
            else:
                object.__setattr__(self, attr, val)

    def __get__(self, obj, objtype):
        """Declare the variable and return the property."""
        self.declare_once()
        return self

    def __set__(self, obj, value):
        """Declare the variable and assign ``value`` to it."""
        self.declare_once()
        self.bsca.addressing.note_assignment(self, value)
        self.bsca.append_code("%s = %s;\n" % (self, value))

    def unpack_cast(self, expr):
        """Cast the unpacked integer bits to the property's type."""
        return expr

    def pack_cast(self, expr):
        """Cast the property's value to integer bits for packing."""
        return expr

    @property
    def declared(self):
        """Test if the state variable is declared."""
        return self.bsca.is_declared(self)

    @property
    def coords_declared(self):
        """Test if the coordinate variables are declared."""
        return self.bsca.coords_declared

    def declare_once(self):
        """Declare the C variable holding the property (once per kernel)."""
        if self.declared:
            return
        self.bsca.append_code("%s %s;\n" % (self.ctype, self.var_name))
        self.bsca.declare(self)

    def __str__(self):
        """Return the variable name to use in mixed expressions."""
        return self.var_name


class IntegerProperty(Property):
    """A non-negative integer with the upper limit ``max_val``."""

    def __init__(self, max_val):
        """Initialize class specific attributes."""
        self.max_val = max_val
        super().__init__()

    def calc_bit_width(self):
        """``int(log2(max_val)) + 1`` bits (``properties.py:304-306``)."""
        return int(math.log2(self.max_val)) + 1


class FloatProperty(Property):
    """A 32-bit float, packed by bit-cast (``properties.py:309-332``)."""

    def calc_bit_width(self):
        """Always 32 bits."""
        return 32

    @cached_property
    def ctype(self):
        """Always ``float``."""
        return "float"

    def unpack_cast(self, expr):
        """Reinterpret the integer bits as a float."""
        return f"(* reinterpret_cast<float*>(&({expr})))"

    def pack_cast(self, expr):
        """Reinterpret the float as integer bits."""
        return f"(* reinterpret_cast<int*>(&({expr})))"


class ContainerProperty(Property):
    """
    The packed word holding all top-level properties of a cell.

    Used by the framework for the main state, each buffer and each
    neighbour's main state and buffer.
    """

    def __init__(self):
        """Initialize the ordered field table."""
        super().__init__()
        self._properties = OrderedDict()
        # Set once any field is read: from then on (in every kernel) the
        # container is declared initialised from memory (the reference
        # keeps the same sticky ``init_val``, ``properties.py:409``).
        self.init_val = None

    @property
    def properties(self):
        """Get inner properties."""
        return self._properties

    def values(self):
        """Iterate over properties, emulating ``dict`` functionality."""
        yield from self._properties.values()

    @property
    def unpacked(self):
        """Test if inner properties are unpacked from memory."""
        return self.bsca.is_unpacked(self)

    def __getitem__(self, key):
        """Get a property by key."""
        return self._properties[key]

    def __setitem__(self, key, val):
        """Set a property by key."""
        self._properties[key] = val
        object.__setattr__(self, key, val)

    def calc_bit_width(self):
        """Sum of the inner properties' bit widths."""
        return sum(p.bit_width for p in self._properties.values())

    @cached_property
    def best_type(self):
        """The smallest word holding every field: up to 128 bits."""
        if self.bit_width <= 64:
            return super().best_type
        if self.bit_width <= 128:
            return _WIDE_TYPE
        raise XenticaException(
            "cell state of %d bits does not fit one 128-bit word"
            % self.bit_width)

    @Property.buf_num.setter
    def buf_num(self, val):
        """Set the memory plane and propagate it to inner properties."""
        self._buf_num = val
        for key in self._properties.keys():
            self[key].buf_num = val

    @Property.nbr_num.setter
    def nbr_num(self, val):
        """Set the neighbour index and propagate it to inner properties."""
        self._nbr_num = val
        for key in self._properties.keys():
            self[key].nbr_num = val

    def __get__(self, obj, objtype):
        """Return the container itself."""
        return self

    def __set__(self, obj, value):
        """Ignore assignment of the container as a whole."""

    def _mem_cell(self):
        """C lvalue of the container's word, from the addressing scheme."""
        return self.bsca.addressing.cell_ref(self)

    def __getattribute__(self, attr):
        """Reading a field loads and unpacks the whole word first."""
        obj = object.__getattribute__(self, attr)
        if isinstance(obj, Property):
            self.init_val = True
            self.declare_once()
            self._unpack_state()
            return obj.__get__(self, type(self))
        return obj

    def __setattr__(self, attr, val):
        """Writing a field declares it and schedules the packed store."""
        try:
            obj = object.__getattribute__(self, attr)
        except AttributeError:
            object.__setattr__(self, attr, val)
        else:
            if isinstance(obj, Property):
                obj.declare_once()
                obj.__set__(self, val)
                self.declare_once()
                self.bsca.deferred_write(self)
                self.bsca.unpack(self)
            else:
                object.__setattr__(self, attr, val)

    def declare_once(self):
        """Declare the word variable (and neighbour access code) once."""
        if self.declared:
            return
        addressing = self.bsca.addressing
        if self._nbr_num >= 0:
            code = addressing.neighbor_declaration(self)
        elif self.init_val is None:
            code = "%s %s;\n" % (self.ctype, self.var_name)
        else:
            code = "%s %s = %s;\n" % (
                self.ctype, self.var_name, self._mem_cell())
        self.bsca.append_code(code)
        self.bsca.declare(self)

    def _unpack_state(self):
        """Unpack every field from the word (``properties.py:496-514``)."""
        if self.unpacked:
            return
        tmp_var = "tmp_%s" % self.var_name
        lines = ["int %s;\n" % tmp_var]
        shift = 0
        for prop in self._properties.values():
            prop.declare_once()
            val = self.var_name
            if shift > 0:
                val += " >> %d" % shift
            lines.append("%s = (%s) & %d;\n" % (
                tmp_var, val, 2 ** prop.bit_width - 1))
            lines.append("%s = %s;\n" % (
                prop.var_name, prop.unpack_cast(tmp_var)))
            shift += prop.bit_width
        self.bsca.append_code("".join(lines))
        self.bsca.unpack(self)

    def deferred_write(self):
        """Pack the fields and store the word (``properties.py:516-541``)."""
        shift = 0
        vals = []
        for prop in self._properties.values():
            prop.declare_once()
            val = "((%s) %s & %d)" % (
                self.ctype, prop.pack_cast(prop.var_name),
                2 ** prop.bit_width - 1)
            if shift > 0:
                val = "(%s << %d)" % (val, shift)
            vals.append(val)
            shift += prop.bit_width
        code = "%s = %s;\n" % (self.var_name, " + ".join(vals))
        code += "%s = %s;\n" % (
            self.bsca.addressing.store_ref(self), self.var_name)
        self.bsca.append_code(code)


class TotalisticRuleProperty(Property):
    """
    A per-cell totalistic rule (``properties.py:544-619``).

    Width is ``2 * (k + 1)`` bits for ``k`` neighbours: birth bits first,
    then survival bits.
    """

    def __init__(self, outer=False):
        """Initialize the rule."""
        self._outer = outer
        super().__init__()

    @property
    def _genome_mask(self):
        """Mask of the meaningful birth bits."""
        return 2 ** (self._num_neighbors + 1) - 2

    def calc_bit_width(self):
        """``(k + 1) * 2`` bits."""
        return (self._num_neighbors + 1) * 2

    def is_sustained(self, num_neighbors):
        """Survival flag for ``num_neighbors`` live neighbours."""
        if not self._outer:
            raise XenticaException(
                "Can not get a sustained flag from the pure totalistic rule.")
        mask = self._genome_mask << (self._num_neighbors + 1)
        return ((self & mask) >> (
            num_neighbors + self._num_neighbors + 1)) & 1

    def is_born(self, num_neighbors):
        """Birth flag for ``num_neighbors`` live neighbours."""
        if not self._outer:
            raise XenticaException(
                "Can not get a born flag from the pure totalistic rule.")
        return ((self & self._genome_mask) >> num_neighbors) & 1

    def __get__(self, obj, objtype):
        """Declare and return the property."""
        self.declare_once()
        return self

    def next_val(self, cur_val, num_neighbors):
        """Next state of a cell in state ``cur_val``."""
        if not self._outer:
            return (self >> (num_neighbors + cur_val)) & 1
        alive_shift = cur_val * (self._num_neighbors + 1)
        return (self >> (num_neighbors + alive_shift)) & 1


class RandomProperty(Property):
    """
    A 16-bit LCG that advances on every access (``properties.py:622-660``).

    ``x.rng`` emits ``rng = ((unsigned)rng * 58321 + 11113) % 65535`` and
    schedules the container's store; ``x.rng.uniform`` is the value that
    access produced, divided by 65535 as a float.
    """

    def __get__(self, obj, objtype):
        """Advance the generator, then return the property."""
        self.declare_once()
        val_int = xmath.int(DeferredExpression(self.var_name))
        self.__set__(self, ((val_int * 58321 + 11113)) % 65535)
        self.bsca.deferred_write(obj)
        return self

    def calc_bit_width(self):
        """Always 16 bits."""
        return 16

    @property
    def uniform(self):
        """The current value as a float in [0, 1]."""
        return xmath.float(DeferredExpression(self.var_name)) / 65535


class XorShiftProperty(Property):
    """
    A 32-bit xorshift generator that advances on every access.

    No counterpart in the reference, whose only generator is the LCG of
    :class:`RandomProperty` (``properties.py:622-660``); the access
    semantics are the same: each ``x.rng`` read emits one advance and
    schedules the container's store, ``x.rng.uniform`` is the value that
    access produced as a float in [0, 1].  The advance is Marsaglia's
    xorshift with shifts (13, 17, 5) (G. Marsaglia, "Xorshift RNGs",
    Journal of Statistical Software 8(14), 2003): period 2**32 - 1 over
    the non-zero states, so seed every cell with a non-zero value.
    """

    def __get__(self, obj, objtype):
        """Advance the generator, then return the property."""
        self.declare_once()
        var = self.var_name
        for c_op, shift in (("<<", 13), (">>", 17), ("<<", 5)):
            self.__set__(self, DeferredExpression(
                "(%s ^ (%s %s %d))" % (var, var, c_op, shift)))
        self.bsca.deferred_write(obj)
        return self

    def calc_bit_width(self):
        """Always 32 bits."""
        return 32

    @property
    def uniform(self):
        """The current value as a float in [0, 1]."""
        return xmath.float(DeferredExpression(self.var_name)) / 4294967295
