"""Formatters (``tests/utils/test_formatters.py`` of the reference)."""
from xentica_amd.utils.formatters import sizeof_fmt


def test_formats():
    assert sizeof_fmt(345) == "345"
    assert sizeof_fmt(34567) == "34.57K"
    assert sizeof_fmt(34567 * (10 ** 12)) == "34.57P"
    assert sizeof_fmt(34567 * (10 ** 21)) == "34.57Y"
