"""The exception raised for every DSL, topology and runtime error.

Mirrors ``xentica/core/exceptions.py:6`` (one exception class for the
whole framework).
"""

__all__ = ['XenticaException', ]


class XenticaException(Exception):
    """Basic Xentica exception."""
