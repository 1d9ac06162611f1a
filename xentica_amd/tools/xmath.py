"""
C math intrinsics usable in DSL expressions (``xentica/tools/xmath.py``).

``min``/``max`` nest C ternaries left to right, ``int`` casts to
``unsigned int``, ``float`` to ``float``; ``popc`` and ``exp`` call the
device intrinsics (``__popc`` is a HIP builtin on gfx950).
"""
from xentica_amd.core.expressions import DeferredExpression


class Xmath:
    """Static helpers returning :class:`DeferredExpression` objects."""

    @staticmethod
    def min(*args):
        """Minimum of the arguments as nested ternaries."""
        expr = str(args[0])
        for arg in args[1:]:
            expr = "(({a}) < ({b})) ? ({a}) : ({b})".format(a=expr, b=arg)
        return DeferredExpression("(%s)" % expr)

    @staticmethod
    def max(*args):
        """Maximum of the arguments as nested ternaries."""
        expr = str(args[0])
        for arg in args[1:]:
            expr = "(({a}) > ({b})) ? ({a}) : ({b})".format(a=expr, b=arg)
        return DeferredExpression("(%s)" % expr)

    @staticmethod
    def float(val):
        """Cast to ``float``."""
        return DeferredExpression("((float) (%s))" % val)

    @staticmethod
    def int(val):
        """Cast to ``unsigned int``."""
        return DeferredExpression("((unsigned int) (%s))" % val)

    @staticmethod
    def popc(val):
        """Population count."""
        return DeferredExpression("(__popc(%s))" % val)

    @staticmethod
    def exp(val):
        """Natural exponent."""
        return DeferredExpression("(exp(%s))" % val)
