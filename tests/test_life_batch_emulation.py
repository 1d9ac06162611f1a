"""
The bit-sliced batch kernel's plan and logic, emulated on the CPU, against
the oracle.

``tests/emulation/life_batch.py`` restates the kernel's per-lane set-up,
row pipeline and buffer offsets in NumPy over the lane descriptors that
``runtime/bits_batch.plan_lanes`` produces; these tests show that every
launch stays in bounds, stores every word of every instance exactly once
and reproduces the oracle for ragged widths (not multiples of 32), heights
shorter than a strip or than the halo, tiles of wide rows, several
instances sharing a wave, and rules of every kind.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.emulation import life_batch as E
from xentica_amd.runtime.bits_batch import (
    plan_lanes, layout, pitch_of, balanced_strips)


def pack_rows(cells, width, height):
    pitch = pitch_of(width)
    padded = np.zeros((height, pitch * 32), dtype=np.uint64)
    padded[:, :width] = cells.reshape(height, width)
    words = np.zeros((height, pitch), dtype=np.uint64)
    for bit in range(32):
        words |= padded[:, bit::32] << np.uint64(bit)
    return words.reshape(-1)


def unpack_rows(words, width, height):
    pitch = pitch_of(width)
    words = words.reshape(height, pitch)
    out = np.zeros((height, pitch * 32), dtype=np.uint8)
    for bit in range(32):
        out[:, bit::32] = (words >> np.uint64(bit)) & np.uint64(1)
    return out[:, :width].reshape(-1)


def _check(sizes, rules, steps, depth, strip_rows, seed=0):
    rng = np.random.default_rng(seed)
    bases, total = layout(sizes)
    src = np.zeros(total, dtype=np.uint64)
    starts = []
    for (width, height), base in zip(sizes, bases):
        cells = rng.integers(0, 2, width * height, dtype=np.uint8)
        starts.append(cells)
        src[base:base + pitch_of(width) * height] = pack_rows(cells, width,
                                                              height)
    insts = np.zeros((2 * len(sizes), 4), dtype=np.int64)
    for idx, ((width, height), base, (born, sustain)) in enumerate(
            zip(sizes, bases, rules)):
        insts[2 * idx] = (base, pitch_of(width), width, height)
        insts[2 * idx + 1] = (born, sustain, 0, 0)
    lanes, wave_rows = plan_lanes(sizes, strip_rows)
    assert np.all((lanes[:, :, 3] > 0).sum(axis=1) <= 62 * 64)
    left = steps
    while left:
        k = min(depth, left)
        src, stores = E.kernel(src, 4 * total, insts, lanes, wave_rows, k)
        assert np.all(stores == 1), "every word stored exactly once"
        left -= k
    for idx, ((width, height), base, (born, sustain)) in enumerate(
            zip(sizes, bases, rules)):
        field = O.Field((width, height), 8, np.uint8, starts[idx])
        O.lifelike(field, O.moore_deltas(2), O.moore_rev(2), born, sustain,
                   steps, colors=False)
        got = unpack_rows(src[base:base + pitch_of(width) * height], width,
                          height)
        assert np.array_equal(got, field.cells), (idx, width, height)


def test_ragged_widths_share_waves():
    sizes = [(32, 5), (33, 7), (45, 9), (63, 4), (64, 13), (70, 1),
             (100, 3), (95, 11)]
    rules = [(8, 12), (72, 12), (1, 256), (511, 0), (0, 511), (0b101000,
             0b1100), (8, 12), (0b111101000, 0b111100000)]
    _check(sizes, rules, steps=11, depth=4, strip_rows=4)


@pytest.mark.parametrize("depth", [1, 3, 8])
def test_depths_and_short_fields(depth):
    sizes = [(40, 2), (96, 17), (37, 3)]
    rules = [(8, 12), (0b1001000, 0b1100), (0b100, 0b10010)]
    _check(sizes, rules, steps=2 * depth + 1, depth=depth, strip_rows=6,
           seed=depth)


def test_wide_rows_split_into_tiles():
    """Rows of more than 62 words take several column tiles; their halo
    lanes reach into the neighbouring tiles and across the row's end."""
    sizes = [(62 * 32 + 17, 6), (4000, 3)]
    rules = [(8, 12), (0b11000, 0b1101)]
    _check(sizes, rules, steps=5, depth=5, strip_rows=3, seed=3)


def test_plan_covers_the_batch_workload():
    """The 256-instance batch: every stored word has exactly one storing
    lane and no wave holds more than 64 lanes (checked on the plan without
    running the kernel)."""
    from examples.workloads import batch_experiments
    sizes = [exp.size for exp in batch_experiments()]
    bases, total = layout(sizes)
    for strip_rows in (16, 39, 48):
        lanes, wave_rows = plan_lanes(sizes, strip_rows)
        count = np.zeros(total, dtype=np.int64)
        active = lanes[:, :, 3] > 0
        for inst, col, row_lo, rows in lanes[active]:
            width, height = sizes[inst]
            assert 0 <= col < pitch_of(width)
            assert 0 <= row_lo and row_lo + rows <= height
            start = bases[inst] + row_lo * pitch_of(width) + col
            count[start:start + rows * pitch_of(width):pitch_of(width)] += 1
        assert np.all(count == 1)
        assert np.all(wave_rows >= lanes[:, :, 3].max(axis=1))


def test_balanced_strips_fill_whole_waves_per_simd():
    """The heights the autotuner adds: the lowest whose plan has at most
    two (three) waves per SIMD of a 1024-SIMD device, so one row less
    would exceed that count."""
    from examples.workloads import batch_experiments
    sizes = [exp.size for exp in batch_experiments()]
    heights = balanced_strips(sizes, 1024)
    assert len(heights) == 2
    for height, per_simd in zip(heights, (3, 2)):
        assert len(plan_lanes(sizes, height)[1]) <= per_simd * 1024
        assert len(plan_lanes(sizes, height - 1)[1]) > per_simd * 1024
    # a batch too small to reach the counts gets no height
    assert balanced_strips([(64, 64)], 1024) == []
