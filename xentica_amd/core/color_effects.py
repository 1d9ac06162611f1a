"""
Colour effects: decorators of a model's ``color()`` method.

``color()`` returns ``(red, green, blue)`` DSL expressions; the decorator
renders them into the absorb kernel and stores the cell's colour as
``col[i] = make_int3(r, g, b)`` (``xentica/core/color_effects.py:66-82``).
:class:`MovingAverage` first scales by ``SMOOTH_FACTOR`` and clamps each
channel to ``[old - FADE_OUT, old + FADE_IN]`` (``color_effects.py:106-130``;
defaults 255, 255 and 1 when the experiment sets none).

Example::

    @color_effects.MovingAverage
    def color(self):
        v = self.main.state * 255
        return (v, v, v)
"""
from xentica_amd.core.variables import Constant
from xentica_amd.core.mixins import BscaDetectorMixin

__all__ = ['ColorEffect', 'MovingAverage', ]


class ColorEffect(BscaDetectorMixin):
    """
    The base colour effect: stores the colour as computed.

    Subclasses put C code transforming ``new_r``, ``new_g`` and ``new_b``
    into ``self.effect`` and then call ``super().__call__()``.
    """

    def __init__(self, func):
        """Wrap the model's colour function."""
        self.func = func
        self.effect = ""
        #: the statements of ``effect`` that read the stored colour (``""``:
        #: the colour written never depends on the colour before it)
        self.memory = ""

    def __call__(self):
        """Emit the colour computation into the current kernel."""
        red, green, blue = self.func(self.bsca)
        code = (
            "int new_r = %s;\n"
            "int new_g = %s;\n"
            "int new_b = %s;\n"
            "%s\n"
            "col[i] = make_int3(new_r, new_g, new_b);\n"
        ) % (red, green, blue, self.effect)
        self.bsca.append_code(code)


class MovingAverage(ColorEffect):
    """
    Per-channel moving average with bounded fade in / fade out.

    Experiment settings: ``fade_in``, ``fade_out`` (largest increase and
    decrease per step) and ``smooth_factor`` (scale applied first).
    """

    def __call__(self):
        """Emit the clamped colour update."""
        bsca = self.bsca
        if not hasattr(bsca, "fade_in"):
            bsca.fade_in = 255
        if not hasattr(bsca, "fade_out"):
            bsca.fade_out = 255
        if not hasattr(bsca, "smooth_factor"):
            bsca.smooth_factor = 1
        bsca.define_constant(Constant("FADE_IN", bsca.fade_in))
        bsca.define_constant(Constant("FADE_OUT", bsca.fade_out))
        bsca.define_constant(Constant("SMOOTH_FACTOR", bsca.smooth_factor))
        clamp = ("new_{c} = max(min(new_{c}, old_col.{x} + FADE_IN), "
                 "old_col.{x} - FADE_OUT);\n")
        if (isinstance(bsca.fade_in, int) and isinstance(bsca.fade_out, int)
                and bsca.fade_in + bsca.fade_out >= 0):
            # lo = old - FADE_OUT <= hi = old + FADE_IN, so the clamp is
            # the median of (new, lo, hi): one v_med3_i32 instead of a
            # min and a max
            clamp = ("new_{c} = max(min(new_{c}, old_col.{x} - FADE_OUT), "
                     "min(max(new_{c}, old_col.{x} - FADE_OUT), "
                     "old_col.{x} + FADE_IN));\n")
        self.memory = ("int3 old_col = col[i];\n"
                       + "".join(clamp.format(c=c, x=x)
                                 for c, x in (("r", "x"), ("g", "y"),
                                              ("b", "z"))))
        self.effect = (
            "new_r *= SMOOTH_FACTOR;\n"
            "new_g *= SMOOTH_FACTOR;\n"
            "new_b *= SMOOTH_FACTOR;\n"
            + self.memory)
        return super().__call__()
