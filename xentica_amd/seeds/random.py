"""
Per-model random streams (``xentica/seeds/random.py``).

:class:`LocalRandom` pairs a stdlib ``random.Random(seed)`` with a NumPy
legacy ``RandomState`` seeded from that stream's first 32 random bits
(``random.py:40-44``), so one experiment word reproduces both streams.
:class:`RandInt` draws one value per cell of the pattern area from the
NumPy stream, or one value from the stdlib stream when ``constant=True``
(``random.py:81-96``).
"""
import random
import functools
import operator

import numpy as np

from xentica_amd.core.expressions import PatternExpression

__all__ = ['LocalRandom', 'RandInt', ]


class LocalRandom:
    """The stdlib and NumPy random streams of one model."""

    def __init__(self, seed=None):
        """Seed both streams from ``seed``."""
        self.standard = random.Random(seed)
        self.numpy = np.random.RandomState(self.standard.getrandbits(32))

    def load(self, rng):
        """Adopt the streams of another :class:`LocalRandom`."""
        self.standard = rng.standard
        self.numpy = rng.numpy


class RandInt(PatternExpression):
    """
    Random integers in ``[min_val, max_val]``.

    :param min_val: lower bound (inclusive)
    :param max_val: upper bound (inclusive)
    :param constant: draw a single value from the stdlib stream
    """

    def __init__(self, min_val, max_val, constant=False):
        """Store the interval."""
        super().__init__()
        self.min_val = min_val
        self.max_val = max_val
        self._constant = constant

    def __get__(self, instance, owner):
        """Draw the values for ``instance`` (a pattern with ``size``)."""
        if hasattr(instance, "size") and not self._constant:
            num_values = functools.reduce(operator.mul, instance.size)
            return instance.random.numpy.randint(
                self.min_val, self.max_val + 1, num_values)
        instance = instance or owner
        return instance.random.standard.randint(self.min_val, self.max_val)
