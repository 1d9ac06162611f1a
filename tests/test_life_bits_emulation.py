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


@pytest.mark.parametrize("words,height,depth,strips", [
    (67, 120, 4, 12),    # one whole tile + 5 columns: 4 segments per wave
    (16, 96, 3, 32),     # no whole tile: 16 columns, 3 segments per wave
    (63, 150, 8, 10),    # 1 column left: 4 segments per wave
    (140, 90, 2, 9),     # two whole tiles + 16 columns
    (80, 77, 5, 7),      # 18 columns, segments of 11 rows (the last 0)
    (84, 60, 2, 20)])    # 22 columns: 2 segments per wave
def test_shared_remainder_waves(words, height, depth, strips):
    """Interior segments of the remainder column share waves; every word is
    stored once, every load of an active lane is in the field, and the
    result is the oracle's."""
    plan = E.make_plan(words, height, depth, strips)
    assert plan["e_hi"] - plan["e_lo"] >= 2 and plan["group"] >= 2
    _check(32 * words, height, depth, 8, 12, depth + 1, strips=strips)
    _check(32 * words, height, depth, 6, 20, depth, strips=strips, seed=2)


def test_wide_remainder_runs_one_segment_per_wave():
    """31 or more columns left over: no sharing (one segment per wave)."""
    plan = E.make_plan(93, 60, 5, 6)
    assert plan["group"] == 1 and plan["e_hi"] == plan["e_lo"]
    _check(32 * 93, 60, 5, 8, 12, 7, strips=6)


def test_plan_of_the_16384_field():
    """16384^2 at depth 8: 8 whole tiles and 16 columns per strip; the
    default plan fills at most three waves per SIMD, and 95 % or more of
    its lanes compute output words (89 % with nine 57-column tiles)."""
    strips = E.auto_strips(512, 16384, 8)
    plan = E.make_plan(512, 16384, 8, strips)
    assert plan["full"] == 8 and plan["rem"] == 16 and plan["group"] == 3
    assert 3000 <= plan["waves"] <= 3 * 1024
    out_lanes = 512 * 16384 / (16384 / strips)   # output word-columns x strips
    assert out_lanes / (64 * plan["waves"]) >= 0.95


def _plan_coverage(words, height, depth, strips):
    """Store count of every word and the loaded row range of every wave's
    active segments under the plan (the kernel's offsets, without the
    row pipeline)."""
    plan = E.make_plan(words, height, depth, strips)
    count = np.zeros((height, words), dtype=np.int32)
    for wave in range(plan["waves"]):
        col_lo, ncols, row_lo, out, seg_lanes, nsegs, exchange = E.wave_tile(
            plan, wave, height, depth)
        assert 1 <= ncols <= E.OUT_LANES and 1 <= out
        assert nsegs * seg_lanes <= 64
        if exchange:
            # the next wave of the workgroup holds the next strip's rows
            assert out >= 2 * depth and (wave + 1) % E.G != 0
            nxt = E.wave_tile(plan, wave + 1, height, depth)
            assert nxt[0] == col_lo and nxt[2] == row_lo + out
        last = out + 2 * depth + E.AHEAD - 1     # rows requested: 0 .. last
        for g in range(nsegs):
            lo = row_lo + g * plan["seg_rows"]
            if nsegs > 1:
                # shared segments: the whole window, rows requested ahead
                # included, lies in the field without wrapping
                assert out == plan["seg_rows"]
                assert lo - depth >= 0 and lo - depth + last <= height - 1
            count[lo:lo + out, col_lo:col_lo + ncols] += 1
    return count


@pytest.mark.parametrize("words", [1, 2, 15, 16, 31, 33, 61, 62, 63, 67, 93,
                                   124, 140, 512])
@pytest.mark.parametrize("height", [1, 7, 18, 37, 90, 361, 4096])
def test_plan_stores_every_word_once(words, height):
    for depth in (1, 4, 8):
        strips_list = {1, 3, min(height, 64), height,
                       E.auto_strips(words, height, depth)}
        for strips in sorted(s for s in strips_list if s <= height):
            count = _plan_coverage(words, height, depth, strips)
            assert np.all(count == 1), (words, height, depth, strips)


@pytest.mark.parametrize("words,height,depth,strips", [
    (62, 200, 4, 8),     # two groups of 100 rows: strips of 25, last 25
    (67, 300, 8, 4),     # one group, 5 columns left in halo segments
    (63, 143, 4, 12),    # groups of 47 and 48 rows
    (124, 170, 4, 8),    # two whole tiles, strips of 22 rows (the last 19)
    (140, 260, 5, 8)])   # two whole tiles + 16 columns, depth 5
def test_exchanging_groups(words, height, depth, strips):
    """Strips of a workgroup hand their last rows over to the next strip
    through LDS instead of computing halo rows: the plan is an exchanging
    one, every word is stored once and the result is the oracle's, for
    B3/S23 and a runtime rule."""
    assert E.make_plan(words, height, depth, strips)["xgroups"] == strips // 4
    _check(32 * words, height, depth, 8, 12, depth + 1, strips=strips)
    _check(32 * words, height, depth, 6, 20, depth, strips=strips, seed=3)


def test_exchange_needs_strips_of_2k_rows():
    """Groups whose strips would be shorter than 2K rows run as halo
    strips."""
    assert E.make_plan(62, 100, 8, 8)["xgroups"] == 0   # hx = 13 < 16
    assert E.make_plan(62, 200, 8, 8)["xgroups"] == 2   # hx = 26
    assert E.make_plan(62, 200, 8, 6)["xgroups"] == 0   # not whole groups
    assert E.make_plan(62, 200, 3, 8)["xgroups"] == 0   # depth below 4
