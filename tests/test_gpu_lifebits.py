"""
The bit-packed Life-like engine against the oracle.

Small and ragged fields are checked step by step against the CPU oracle
(every step's CRC-32), over all temporal depths, the fast B3/S23 path and
runtime rules.  At the full 16384x16384 size the engine is checked against
the oracle for the first steps and against the plain GPU engine (the
reference's structure) every 50 steps for 500 steps, through the device
CRC.
"""
import binascii
import zlib

import numpy as np
import pytest

from oracle import oracle as O
from xentica_amd.runtime import hip, kernels
from xentica_amd.runtime.lifebits import LifeBitsField
from xentica_amd.tools.rules import LifeLike

pytestmark = pytest.mark.gpu


def oracle_hashes(cells, width, height, born, sustain, steps):
    field = O.Field((width, height), 8, np.uint8, cells)
    return O.lifelike(field, O.moore_deltas(2), O.moore_rev(2), born, sustain,
                      steps, colors=False, hashes=True), field


def rule_masks(golly):
    rule = LifeLike.golly2int(golly)
    return rule & 0x1FF, (rule >> 9) & 0x1FF


@pytest.mark.parametrize("width,height", [
    (32, 3), (32, 40), (64, 64), (96, 17), (2016, 37), (1984, 70),
    (4096, 150), (128, 1)])
@pytest.mark.parametrize("depth", [1, 3, 4, 8])
def test_shapes_and_depths_b3s23(width, height, depth):
    rng = np.random.default_rng(width * 1000 + height + depth)
    cells = rng.integers(0, 2, width * height, dtype=np.uint8)
    steps = 2 * depth + 3
    want, field = oracle_hashes(cells, width, height, 8, 12, steps)
    life = LifeBitsField(width, height, 8, 12, depth=1)
    life.set_cells(cells)
    for step in range(steps):
        life.launch(depth=1)
        assert life.crc32() == want[step], "depth-1 step %d" % (step + 1)
    life = LifeBitsField(width, height, 8, 12, depth=depth)
    life.set_cells(cells)
    life.run(steps)
    life.synchronize()
    assert np.array_equal(life.get_cells(), field.cells)


@pytest.mark.parametrize("golly", [
    "B3/S23", "B36/S23", "B35678/S5678", "B2/S", "B0/S8", "B012345678/S",
    "B/S012345678", "B1357/S1357", "B3678/S34678"])
def test_rules(golly):
    born, sustain = rule_masks(golly)
    width, height = 320, 45
    rng = np.random.default_rng(7)
    cells = rng.integers(0, 2, width * height, dtype=np.uint8)
    want, field = oracle_hashes(cells, width, height, born, sustain, 12)
    for depth in (1, 5):
        life = LifeBitsField(width, height, born, sustain, depth=depth)
        life.set_cells(cells)
        got = []
        for _ in range(12 // depth + 1):
            life.launch(depth=depth)
        life2 = LifeBitsField(width, height, born, sustain, depth=depth)
        life2.set_cells(cells)
        life2.run(12)
        life2.synchronize()
        assert life2.crc32() == want[11], golly
    life = LifeBitsField(width, height, born, sustain, depth=1)
    life.set_cells(cells)
    for step in range(12):
        life.launch(depth=1)
        got.append(life.crc32())
    assert got == list(want)


def test_pack_unpack_and_device_crc():
    rng = np.random.default_rng(3)
    n = 32 * 12345
    cells = rng.integers(0, 2, n, dtype=np.uint8)
    life = LifeBitsField(32 * 12345, 1)
    life.set_cells(cells)
    assert np.array_equal(life.get_cells(), cells)
    assert life.crc32() == zlib.crc32(cells.tobytes())
    bits = np.packbits(cells.reshape(-1, 8)[:, ::-1], axis=1).reshape(-1)
    assert np.array_equal(life.get_bits().view(np.uint8), bits)
    data = rng.integers(0, 256, 1000003, dtype=np.uint8)
    buf = hip.Buffer(data.size)
    buf.upload(data)
    assert kernels.crc32_u8(buf.ptr, data.size) == binascii.crc32(data)


def test_life16k_first_steps_vs_oracle():
    """16384x16384, seed 42: the first 2 steps equal the oracle's."""
    from examples.workloads import Life16kExperiment
    from xentica_amd.seeds.random import LocalRandom
    size = Life16kExperiment.size
    n = size[0] * size[1]
    cells = LocalRandom(Life16kExperiment.word).numpy.randint(0, 2, n)
    cells = cells.astype(np.uint8)
    life = LifeBitsField(size[0], size[1], 8, 12, depth=2)
    life.set_cells(cells)
    want, field = oracle_hashes(cells, size[0], size[1], 8, 12, 2)
    life.run(2)
    life.synchronize()
    assert life.crc32() == want[1]


def test_life16k_vs_plain_engine():
    """500 steps at 16384x16384: bit-packed and plain engines agree."""
    from examples.game_of_life import GameOfLife
    from examples.workloads import Life16kExperiment
    class PlainLife16k(Life16kExperiment):
        engine = "plain"
    model = GameOfLife(PlainLife16k)
    cells = model.gpu.arrays.cells.get()[:model.cells_num]
    life = LifeBitsField(16384, 16384, 8, 12, depth=4)
    life.set_cells(cells)
    for block in range(10):
        for _ in range(50):
            model.step()
        life.run(50)
        life.synchronize()
        plain = kernels.crc32_u8(model.gpu.arrays.cells.device_ptr(),
                                 model.cells_num)
        assert life.crc32() == plain, "after %d steps" % (50 * (block + 1))


@pytest.mark.parametrize("strips", [0, 455])
def test_life16k_2000_steps_vs_plain_engine(strips):
    """
    The headline configuration over the whole 2000 steps, at the launch
    shape the benchmark times (depth 8; the library's default strips, 368
    on 256 CUs: eight whole 62-column tiles and a third of a shared
    remainder wave per strip, three waves per SIMD; and 455 strips, about
    3.7 waves per SIMD): the first 64 steps one launch each with the CRC of
    every step, the rest in 8-step launches with the CRC after each, all
    equal to the plain engine's (the reference's structure).
    """
    from examples.game_of_life import GameOfLife
    from examples.workloads import Life16kExperiment
    class PlainLife16k(Life16kExperiment):
        engine = "plain"
    model = GameOfLife(PlainLife16k)
    cells = model.gpu.arrays.cells.get()[:model.cells_num]
    life = LifeBitsField(16384, 16384, 8, 12, depth=8, strips=strips)
    life.set_cells(cells)
    print("Life 16384^2, depth 8, %d strips, 2000 steps" % strips, flush=True)
    done = 0
    while done < 2000:
        count = 1 if done < 64 else 8
        for _ in range(count):
            model.step()
        life.run(count)
        life.synchronize()
        done += count
        plain = kernels.crc32_u8(model.gpu.arrays.cells.device_ptr(),
                                 model.cells_num)
        assert life.crc32() == plain, "after %d steps" % done


@pytest.mark.parametrize("words,height,depth,strips", [
    (67, 120, 4, 12), (16, 96, 3, 32), (63, 150, 8, 10), (140, 90, 2, 9),
    (80, 77, 5, 7), (130, 400, 8, 0), (16, 512, 8, 0)])
def test_shared_remainder_waves_vs_oracle(words, height, depth, strips):
    """Fields whose leftover word columns run in shared waves (interior row
    segments of the remainder column packed several to a wave) against
    the oracle, for B3/S23 and a runtime rule, over two launches."""
    width = 32 * words
    rng = np.random.default_rng(words * 7919 + height)
    cells = rng.integers(0, 2, width * height, dtype=np.uint8)
    for born, sustain in ((8, 12), rule_masks("B36/S125")):
        steps = 2 * depth
        want, field = oracle_hashes(cells, width, height, born, sustain, steps)
        life = LifeBitsField(width, height, born, sustain, depth=depth,
                             strips=strips)
        life.set_cells(cells)
        life.run(depth)
        life.synchronize()
        assert life.crc32() == want[depth - 1], (born, sustain, "launch 1")
        life.run(depth)
        life.synchronize()
        assert np.array_equal(life.get_cells(), field.cells), (born, sustain)


@pytest.mark.parametrize("words,height,depth,strips", [
    (62, 200, 4, 8), (67, 300, 8, 4), (63, 143, 4, 12), (124, 170, 4, 8),
    (140, 260, 5, 8), (512, 2048, 8, 16), (512, 1100, 7, 20)])
def test_exchanging_groups_vs_oracle(words, height, depth, strips):
    """Plans whose strips hand their last rows over to the next strip of
    the workgroup through LDS (no halo rows but in each group's last
    strip), against the oracle over two launches, B3/S23 and a runtime
    rule."""
    width = 32 * words
    rng = np.random.default_rng(words * 104729 + height + depth)
    cells = rng.integers(0, 2, width * height, dtype=np.uint8)
    for born, sustain in ((8, 12), rule_masks("B36/S125")):
        steps = 2 * depth
        want, field = oracle_hashes(cells, width, height, born, sustain, steps)
        life = LifeBitsField(width, height, born, sustain, depth=depth,
                             strips=strips)
        life.set_cells(cells)
        life.run(depth)
        life.synchronize()
        assert life.crc32() == want[depth - 1], (born, sustain, "launch 1")
        life.run(depth)
        life.synchronize()
        assert np.array_equal(life.get_cells(), field.cells), (born, sustain)
