"""
Local C variables and ``#define`` constants of the DSL.

Behaviour follows ``xentica/core/variables.py``:

* :class:`Constant` renders ``#define NAME value`` (``variables.py:38-71``);
* :class:`IntegerVariable` is an ``unsigned int``, :class:`FloatVariable` a
  ``float`` (``variables.py:154-173``); a variable is declared with its
  initial value on first use and an assignment through a class-level
  variable emits ``name = value;`` (``variables.py:117-142``);
* the C name of a variable is the Python name it is bound to in the user
  code that first uses it, else the name of the model class attribute
  holding it, else the fallback name (``variables.py:98-115``).
"""
import sys

from xentica_amd.core.mixins import BscaDetectorMixin
from xentica_amd.core.exceptions import XenticaException
from xentica_amd.core.expressions import DeferredExpression

__all__ = ['Constant', 'Variable', 'IntegerVariable', 'FloatVariable', ]

#: Frames of these modules are DSL machinery, never the user's code.
_MACHINERY = frozenset((
    'xentica_amd.core.expressions',
    'xentica_amd.core.variables',
    'xentica_amd.core.properties',
    'xentica_amd.core.parameters',
    'xentica_amd.core.mixins',
    'xentica_amd.core.base',
    'functools',
))
#: Local names that never name a DSL variable (``variables.py:107``).
_BAD_NAMES = ("self_var", "obj", "cls")


def _user_frame():
    """Return the innermost frame that is not DSL machinery."""
    frame = sys._getframe(1)  # pylint: disable=protected-access
    while frame is not None:
        if frame.f_globals.get('__name__') not in _MACHINERY:
            return frame
        frame = frame.f_back
    return None


class Constant(BscaDetectorMixin):
    """
    A ``#define`` pattern.

    Feed it to ``CellularAutomaton.define_constant()`` to have it emitted
    at the top of the generated source.
    """

    def __init__(self, name, value):
        """Store the name and value of the define."""
        self._name = name
        self._value = value
        self.base_class = Constant

    def get_define_code(self):
        """Get the C code for ``#define``."""
        return "#define %s %s\n" % (self._name, str(self._value))

    @property
    def name(self):
        """Get the name of the constant."""
        return self._name

    @property
    def value(self):
        """Get the value of the constant."""
        return self._value


class Variable(DeferredExpression, BscaDetectorMixin):
    """
    The base class for all variables.

    :param val: The initial value for the variable.
    :param name: Fallback name to declare the variable with.
    """

    #: C type to use in the declaration.
    var_type = "unsigned int"

    def __init__(self, val=None, name="var"):
        """Initialize base class features."""
        super().__init__()
        self.fallback_name = name
        self.base_class = Variable
        if val is None:
            raise XenticaException("Variable should have an initial value.")
        self._init_val = DeferredExpression(str(val))

    def _resolve_name(self):
        from xentica_amd.core.base import CellularAutomaton
        frame = _user_frame()
        model = None
        try:
            if frame is not None:
                for key, val in frame.f_locals.items():
                    if isinstance(val, CellularAutomaton):
                        model = val
                    if (val is self and key not in _BAD_NAMES
                            and key.isidentifier()):
                        return key
        finally:
            del frame
        if model is not None and self.fallback_name == "var":
            for key, val in type(model).__dict__.items():
                if val is self and key not in _BAD_NAMES:
                    return key
        return self.fallback_name

    @property
    def var_name(self):
        """Get the C name of the variable (resolved once)."""
        name = self.__dict__.get('_var_name')
        if name is None:
            name = self._resolve_name()
            self.__dict__['_var_name'] = name
        return name

    def declare_once(self):
        """Declare the variable and assign the initial value to it."""
        bsca = self.bsca
        if not bsca.is_declared(self):
            bsca.append_code("%s %s = %s;\n" % (
                self.var_type, self.var_name, self._init_val))
            bsca.declare(self)

    def __str__(self):
        """Return the variable name to use in mixed expressions."""
        return self.var_name

    def __get__(self, obj, objtype):
        """Declare the variable on first use as a class attribute."""
        self.declare_once()
        return self

    def __set__(self, obj, value):
        """Assign a new value to a class-level variable."""
        self.declare_once()
        if str(self.var_name) == str(value):
            return
        self.bsca.append_code("%s = %s;\n" % (self.var_name, value))

    @property
    def code(self):
        """Get the variable name as a C code."""
        return self.var_name

    @code.setter
    def code(self, val):
        """Ignore code changes: a variable always renders as its name."""


class IntegerVariable(Variable):
    """The variable intended to hold a positive integer."""

    var_type = "unsigned int"

    def __init__(self, val="0", **kwargs):
        """Initialize a variable with the default value."""
        super().__init__(val, **kwargs)


class FloatVariable(Variable):
    """The variable intended to hold a 32-bit float."""

    var_type = "float"

    def __init__(self, val="0.0f", **kwargs):
        """Initialize a variable with the default value."""
        super().__init__(val, **kwargs)
