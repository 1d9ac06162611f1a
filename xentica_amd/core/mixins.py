"""
Mixins shared by the DSL classes.

:class:`BscaDetectorMixin` gives DSL objects the model they emit code
into (reference: ``xentica/core/mixins.py:10-53``); :class:`DimensionsMixin`
validates dimensionality (``mixins.py:59-101``).
"""
import inspect

from xentica_amd.core import _context
from xentica_amd.core.exceptions import XenticaException

__all__ = ['BscaDetectorMixin', 'DimensionsMixin', ]


def _model_class():
    from xentica_amd.core.base import CellularAutomaton
    return CellularAutomaton


class BscaDetectorMixin:
    """Find the :class:`CellularAutomaton` instance an object belongs to."""

    @property
    def bsca(self):
        """
        Get the model that code is being generated for.

        During a translation this is the model on the translation stack.
        Otherwise the calling frames are scanned for the nearest model
        instance, as ``xentica/core/mixins.py:23-40`` does.
        """
        model = _context.current()
        if model is not None:
            return model
        ca_class = _model_class()
        frame = inspect.currentframe()
        try:
            while frame is not None:
                for local in frame.f_locals.values():
                    if hasattr(local, "__get__"):
                        continue
                    if isinstance(local, ca_class):
                        return local
                frame = frame.f_back
        finally:
            del frame
        raise XenticaException("BSCA not detected")


class DimensionsMixin:
    """Add a validated ``dimensions`` attribute to a class."""

    #: Supported dimensionalities, set by each subclass.
    supported_dimensions = []

    def __init__(self):
        """Initialize private variables."""
        self._dimensions = None

    def allowed_dimension(self, num_dim):
        """Test if a particular dimensionality is allowed."""
        return num_dim in self.supported_dimensions

    @property
    def dimensions(self):
        """Get a number of dimensions."""
        return self._dimensions

    @dimensions.setter
    def dimensions(self, num_dim):
        """Set a number of dimensions and check if it is allowed."""
        if not self.allowed_dimension(num_dim):
            msg = "%d-D %s is not supported." % (
                num_dim, self.__class__.__name__)
            raise XenticaException(msg)
        self._dimensions = num_dim
