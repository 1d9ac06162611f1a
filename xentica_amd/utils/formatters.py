"""
Number formatting for benchmark output.

The behaviour is what ``tests/utils/test_formatters.py`` of the reference
asserts: values below a thousand print unchanged, larger ones as a
mantissa with two decimals and a decimal SI prefix (K up to Y, Y also for
everything beyond it).
"""

__all__ = ['sizeof_fmt', ]

_PREFIXES = ("K", "M", "G", "T", "P", "E", "Z", "Y")


def sizeof_fmt(num, suffix=''):
    """Format ``num`` with a decimal SI prefix: ``34567`` -> ``"34.57K"``."""
    if abs(num) < 1000:
        return "%s%s" % (num, suffix)
    # the largest power of 1000 that keeps the mantissa at or above 1
    power = 1
    while power < len(_PREFIXES) and abs(num) >= 1000 ** (power + 1):
        power += 1
    return "%.2f%s%s" % (num / 1000 ** power, _PREFIXES[power - 1], suffix)
