"""
Colour helpers for DSL ``color()`` methods (``xentica/tools/color.py``).

:func:`hsv2rgb` emits the branch-free HSV->RGB conversion of
``color.py:8-65`` (selector products instead of branches) when given DSL
expressions, and evaluates it in Python for plain numbers.
:class:`GenomeColor` maps genomes to colours by set-bit positions or by a
modulus (``color.py:68-125``).
"""
from xentica_amd.core.expressions import DeferredExpression
from xentica_amd.core.variables import IntegerVariable, FloatVariable
from xentica_amd.tools import xmath

__all__ = ['hsv2rgb', 'GenomeColor', ]


def hsv2rgb(hue, saturation, value):
    """
    Convert an HSV colour to RGB.

    :param hue: hue in [0, 1]
    :param saturation: saturation in [0, 1]
    :param value: brightness in [0, 1]
    :returns: tuple ``(red, green, blue)``
    """
    if isinstance(hue, DeferredExpression):
        hue_f = FloatVariable()
        hue_f += hue * 6
        hue_i = IntegerVariable()
        hue_i += xmath.int(hue_f)
        hue_f -= hue_i
        sat = FloatVariable()
        sat += saturation
        val = FloatVariable()
        val += value
        grad_p = FloatVariable()
        grad_p += val * (1 - sat)
        grad_q = FloatVariable()
        grad_q += val * (1 - sat * hue_f)
        grad_t = FloatVariable()
        grad_t += val * (1 - sat * (1 - hue_f))
    else:
        hue_f = hue * 6
        hue_i = int(hue_f)
        hue_f -= hue_i
        sat = saturation
        val = value
        grad_p = val * (1 - sat)
        grad_q = val * (1 - sat * hue_f)
        grad_t = val * (1 - sat * (1 - hue_f))

    red = ((hue_i == 0) | (hue_i >= 5)) * val
    red = red + (hue_i == 1) * grad_q
    red = red + ((hue_i == 2) | (hue_i == 3)) * grad_p
    red = red + (hue_i == 4) * grad_t

    green = ((hue_i == 0) | (hue_i >= 6)) * grad_t
    green = green + ((hue_i == 1) | (hue_i == 2)) * val
    green = green + (hue_i == 3) * grad_q
    green = green + ((hue_i == 4) | (hue_i == 5)) * grad_p

    blue = ((hue_i == 0) | (hue_i == 1) | (hue_i >= 6)) * grad_p
    blue = blue + (hue_i == 2) * grad_t
    blue = blue + ((hue_i == 3) | (hue_i == 4)) * val
    blue = blue + (hue_i == 5) * grad_q

    return (red, green, blue)


class GenomeColor:
    """Genome -> RGB colour mappings."""

    @staticmethod
    def positional(genome, num_genes):
        """Blend the hues of the positions of set genes, normalised to max."""
        red = FloatVariable()
        green = FloatVariable()
        blue = FloatVariable()
        for i in range(num_genes):
            d_red, d_green, d_blue = hsv2rgb(i / num_genes, 1, 1 / num_genes)
            val = (genome >> i) & 1
            red += d_red * val
            green += d_green * val
            blue += d_blue * val
        maxval = FloatVariable()
        maxval += xmath.max(red, green, blue)
        red /= maxval
        green /= maxval
        blue /= maxval
        return (red, green, blue)

    @staticmethod
    def modular(genome, divider):
        """Use ``genome % divider`` as a fully saturated hue."""
        return hsv2rgb(genome % divider / xmath.float(divider), 1, 1)
