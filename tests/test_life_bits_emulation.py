"""
The bit-packed Life kernel's logic, emulated on the CPU, against the oracle.

``tests/emulation/life_bits.py`` restates the kernel's loops and lane
operations in NumPy; these tests show that every launch terminates, stays
in bounds, writes every output word once and reproduces the oracle, for
the shapes and rules the GPU tests use.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.emulation import life_bits as E


def _check(width, height, depth, born, sustain, steps, strips=0, seed=0):
    rng = np.random.default_rng(seed)
    cells = rng.integers(0, 2, width * height, dtype=np.uint8)
    got = E.run(cells, width, height, depth, born, sustain, steps, strips)
    field = O.Field((width, height), 8, np.uint8, cells)
    O.lifelike(field, O.moore_deltas(2), O.moore_rev(2), born, sustain, steps,
               colors=False)
    assert np.array_equal(got, field.cells)


@pytest.mark.parametrize("width,height,depth", [
    (32, 3, 1), (32, 40, 3), (64, 64, 4), (96, 17, 8), (2016, 37, 4),
    (128, 1, 8), (4096, 9, 2)])
def test_b3s23_shapes(width, height, depth):
    _check(width, height, depth, 8, 12, 2 * depth + 3)


@pytest.mark.parametrize("born,sustain", [
    (72, 12), (0b111101000, 0b111100000), (1, 256), (511, 0), (0, 511)])
def test_runtime_rules(born, sustain):
    _check(320, 45, 5, born, sustain, 12)


def test_explicit_strips():
    _check(2016, 50, 4, 8, 12, 9, strips=7)


@pytest.mark.parametrize("height,strips", [(30, 10), (63, 9), (64, 8), (72, 8),
                                           (75, 5), (77, 7)])
def test_strip_heights_around_depth(height, strips):
    """Strips of 3, 7, 8, 9, 15 and 11 rows at depth 8: prologue longer
    than the strip, no steady-state iteration, steady state plus a drain
    of every length modulo four."""
    _check(64, height, 8, 8, 12, 8, strips=strips)
    _check(64, height, 8, 6, 20, 8, strips=strips, seed=1)
