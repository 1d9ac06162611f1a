"""
Pin the CPU oracle before trusting it.

Every integer-only golden vector the reference's tests hold is reproduced
here by the oracle (``oracle/oracle.py`` + ``oracle/bsca_oracle.c``);
Life is also checked against an independent NumPy roll-sum and against
closed forms (blinker, glider, block), ShiftingSands and NoiseTV against
their closed forms, the CRC against zlib, and the 3-D Moore offset list
against a brute-force neighbour count.
"""
import zlib

import numpy as np
import pytest

from oracle import oracle as O


def test_crc32_matches_zlib():
    rng = np.random.default_rng(1)
    for size in (0, 1, 7, 4096, 100003):
        data = rng.integers(0, 256, size, dtype=np.uint8)
        assert O.crc32(data) == zlib.crc32(data.tobytes())


def test_random_sequences():
    """``tests/seeds/test_random.py:31-45``."""
    std, _ = O.local_random("test")
    assert [std.randint(11, 23) for _ in range(10)] == \
        [22, 15, 21, 23, 14, 14, 11, 20, 17, 23]
    _, nrs = O.local_random("test")
    assert list(nrs.randint(11, 23, (10,))) == \
        [19, 16, 22, 19, 15, 22, 20, 20, 13, 15]


def test_seed_crc():
    """PrimordialSoup with LocalRandom("test"), ``test_patterns.py:155-164``."""
    rng = O.local_random("test")
    vals, _ = O.primordial_soup((100, 100), {"state": O.randint(0, 1)}, rng)
    cells = np.zeros(10000, np.int32)
    cells[:] = vals["state"]
    assert zlib.crc32(cells.tobytes()) == 1648433356


def test_golly():
    """``tests/tools/test_rules.py:14``."""
    assert O.golly2int("B3/S23") == 6152


def _life(field, steps, **kw):
    O.lifelike(field, O.moore_deltas(2), O.moore_rev(2), 8, 12, steps, **kw)


def test_game_of_life_golden():
    """GameOfLife/GOLExperiment, 1000 steps: ``test_base.py:28-41``."""
    field = O.gol_experiment_field()
    _life(field, 1000)
    assert O.crc32(field.cells) == 2981695958


def test_static_border_golden():
    """GameOfLifeStatic: ``test_base.py:138-147``."""
    field = O.gol_experiment_field()
    _life(field, 1000, border_static=True, border_value=1)
    assert O.crc32(field.cells) == 1098273940


def test_interactive_rule_golden():
    """LifelikeCA/DiamoebaExperiment, rule switch at t=23: ``test_base.py:218-225``."""
    field = O.gol_experiment_field()
    rule = O.golly2int("B35678/S5678")
    O.lifelike(field, O.moore_deltas(2), O.moore_rev(2), rule, rule >> 9, 23)
    rule = O.golly2int("B3/S23")
    O.lifelike(field, O.moore_deltas(2), O.moore_rev(2), rule, rule >> 9, 977)
    assert O.crc32(field.cells) == 4016906756


def test_6d_von_neumann_golden():
    """GameOfLife6D on (64, 36, 3, 3, 3, 3): ``test_base.py:158-168``."""
    field = O.soup_field((64, 36, 3, 3, 3, 3), "XENTICA IS YOUR GODDESS",
                         {"state": O.randint(0, 1)}, [1], ["state"], 12,
                         np.uint8)
    O.lifelike(field, O.von_neumann_deltas(6), O.von_neumann_rev(6), 8, 12,
               1000)
    assert O.crc32(field.cells) == 2742543959


def test_color_life_golden():
    """GameOfLifeColor/GOLExperimentColor: ``test_base.py:149-156``."""
    field = O.soup_field(
        (640, 360), "OBEY XENTICA",
        {"state": O.randint(0, 1), "red": O.randint(0, 255),
         "green": O.randint(0, 255), "blue": O.randint(0, 255)},
        [1, 8, 8, 8], ["state", "red", "green", "blue"], 8, np.uint32)
    O.color_life(field, 1000)
    assert O.crc32(field.cells) == 532957133


def _sands_field():
    rng = O.local_random("LET IT SHIFT")
    idx, vals = O.big_bang((640, 360), (320, 180), (100, 100),
                           {"state": O.randint(0, 1)}, rng)
    state = np.zeros(640 * 360, np.uint8)
    state[idx] = vals["state"]
    return O.Field((640, 360), 8, np.uint8, state)


def test_shifting_sands_golden_and_closed_form():
    """ShiftingSands: ``test_base.py:182-189``; also an exact roll."""
    field = _sands_field()
    start = field.cells.copy().reshape(360, 640)
    O.shifting_sands(field, 1000)
    assert O.crc32(field.cells) == 1639048270
    rolled = np.roll(np.roll(start, -1000, axis=0), -1000, axis=1)
    assert np.array_equal(field.cells.reshape(360, 640), rolled)


def test_noisetv_golden_and_closed_form():
    """NoiseTV: ``test_properties.py:64-71``; also the LCG twice per step."""
    field = O.soup_field((640, 360), "PSHSHSHSH!",
                         {"rng": O.randint(0, 65535)}, [16], ["rng"], 8,
                         np.uint16)
    start = field.cells.astype(np.int64)
    O.noisetv(field, 100)
    assert O.crc32(field.cells) == 2773894957
    r = start
    for _ in range(200):
        r = (r * 58321 + 11113) % 65535
    assert np.array_equal(field.cells, r.astype(np.uint16))


def test_life_against_numpy_rollsum():
    rng = np.random.default_rng(5)
    w0, w1 = 37, 23
    state = rng.integers(0, 2, (w1, w0), dtype=np.uint8)
    field = O.Field((w0, w1), 8, np.uint8, state.reshape(-1))
    for _ in range(25):
        _life(field, 1)
        state = O.numpy_life_step(state)
        assert np.array_equal(field.cells.reshape(w1, w0), state)


def test_life_closed_forms():
    w = 16
    # blinker: period 2
    state = np.zeros((w, w), np.uint8)
    state[5, 4:7] = 1
    field = O.Field((w, w), 8, np.uint8, state.reshape(-1))
    _life(field, 1)
    vertical = np.zeros((w, w), np.uint8)
    vertical[4:7, 5] = 1
    assert np.array_equal(field.cells.reshape(w, w), vertical)
    _life(field, 1)
    assert np.array_equal(field.cells.reshape(w, w), state)
    # glider: moves (+1, +1) every 4 steps on the torus
    glider = np.zeros((w, w), np.uint8)
    glider[1, 2] = glider[2, 3] = glider[3, 1] = glider[3, 2] = glider[3, 3] = 1
    field = O.Field((w, w), 8, np.uint8, glider.reshape(-1))
    _life(field, 4 * w)
    assert np.array_equal(field.cells.reshape(w, w), glider)
    _life(field, 4)
    assert np.array_equal(field.cells.reshape(w, w),
                          np.roll(np.roll(glider, 1, 0), 1, 1))
    # block: still life, also under a dead static border at the corner
    block = np.zeros((w, w), np.uint8)
    block[0:2, 0:2] = 1
    field = O.Field((w, w), 8, np.uint8, block.reshape(-1))
    _life(field, 3, border_static=True, border_value=0)
    assert np.array_equal(field.cells.reshape(w, w), block)


def test_moore_offsets_read_from_reference():
    """
    ``neighborhood.py:165-167`` filters by ``!= (0, 0)``: in 2-D that
    removes the centre; in 3-D it removes nothing, so the centre (index
    13) stays and (1, 1, 1) falls off the end of the 26-entry list.
    """
    d2 = O.moore_deltas(2)
    assert len(d2) == 8 and (0, 0) not in d2
    d3 = O.moore_deltas(3)
    assert len(d3) == 26
    assert d3[13] == (0, 0, 0)
    assert (1, 1, 1) not in d3
    d1 = O.moore_deltas(1)
    assert d1 == [(-1,), (0,)]


def test_3d_lifelike_brute_force():
    """3-D B5/S45 with the reference's offset list, against direct counting."""
    rng = np.random.default_rng(3)
    size = (6, 5, 7)
    state = (rng.random(size[::-1]) < 0.3).astype(np.uint8)  # [x2, x1, x0]
    field = O.Field(size, 26, np.uint8, state.reshape(-1))
    born = (1 << 5)
    sustain = (1 << 4) | (1 << 5)
    for _ in range(4):
        O.lifelike(field, O.moore_deltas(3), O.moore_rev(3), born, sustain, 1)
        count = np.zeros_like(state, dtype=np.int64)
        for d0, d1, d2 in O.moore_deltas(3):
            count += np.roll(state, (-d2, -d1, -d0), axis=(0, 1, 2))
        state = (((born >> count) & 1) | (((sustain >> count) & 1) & state))
        state = state.astype(np.uint8)
        assert np.array_equal(field.cells.reshape(size[::-1]), state)


@pytest.mark.parametrize("threads", [1, 3])
def test_oracle_deterministic_across_threads(threads):
    field = O.gol_experiment_field()
    hashes = O.lifelike(field, O.moore_deltas(2), O.moore_rev(2), 8, 12, 50,
                        threads=threads, hashes=True)
    ref = O.gol_experiment_field()
    want = O.lifelike(ref, O.moore_deltas(2), O.moore_rev(2), 8, 12, 50,
                      threads=8, hashes=True)
    assert np.array_equal(hashes, want)
