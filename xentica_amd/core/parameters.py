"""
Meta-parameters of a model (``xentica/core/parameters.py``).

A non-interactive parameter is inlined as a Python literal and also
registered as a ``#define`` (``parameters.py:126-141``).  An interactive
parameter becomes a kernel argument whose value is read again at every
step (``parameters.py:120-124``, ``base.py:717,722``).  The C and NumPy
types follow the Python type of the value (``parameters.py:78-118``).
"""
import numpy as np

from xentica_amd.core.variables import Constant
from xentica_amd.core.mixins import BscaDetectorMixin
from xentica_amd.core.expressions import DeferredExpression

__all__ = ['Parameter', ]

_CTYPES = {int: 'int', float: 'float', bool: 'bool'}
_DTYPES = {int: np.int32, float: np.float32, bool: bool}


class Parameter(BscaDetectorMixin):
    """
    A model meta-parameter.

    :param default: Value used when the experiment does not set one.
    :param interactive: ``True`` if the value may change between steps.
    """

    def __init__(self, default=0, interactive=False):
        """Initialize the parameter."""
        self._value = default
        self._interactive = interactive
        self._declared = False
        self._name = "param" + str(id(self))

    @property
    def value(self):
        """Get the parameter's value directly."""
        return self._value

    @property
    def interactive(self):
        """Test if the parameter is passed as a kernel argument."""
        return self._interactive

    @property
    def name(self):
        """Get the parameter's name."""
        return self._name

    @name.setter
    def name(self, val):
        """Set the parameter's name."""
        self._name = val

    @property
    def ctype(self):
        """Get the parameter's C type."""
        return _CTYPES.get(type(self._value), 'int32')

    @property
    def dtype(self):
        """Get the parameter's NumPy type."""
        return _DTYPES.get(type(self._value), np.int32)

    def _declare_once(self):
        """Declare the parameter when it's mentioned."""
        if self._interactive:
            if not self.bsca.is_parameter(self.name):
                self.bsca.define_parameter(self)
            return
        if self._declared:
            return
        self._declared = True
        self.bsca.define_constant(Constant(self.name, self._value))

    def __get__(self, obj, objtype):
        """Return the C name (interactive) or the literal value."""
        self._declare_once()
        if self._interactive:
            return DeferredExpression(self._name)
        return self._value

    def __set__(self, obj, value):
        """
        Change the value (read by the next step if interactive).

        The value is class-wide, as in the reference; a model instance also
        remembers its own value so that batched runs can give every
        instance its own parameters.
        """
        self._value = value
        if obj is not None and hasattr(obj, "__dict__"):
            obj.__dict__.setdefault("_param_values", {})[self._name] = value
