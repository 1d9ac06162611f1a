"""
3-D Life-like engines (B5/S45, the reference's 26-offset Moore list) against
the oracle, the DSL plain engine and each other.
"""
import numpy as np
import pytest

from oracle import oracle as O
from examples.workloads import LifeLike3D, Life3DExperiment, life3d_cells
from xentica_amd.runtime.lifebits import Life3DField, Life3DBytes

pytestmark = pytest.mark.gpu

B5, S45 = 1 << 5, (1 << 4) | (1 << 5)


def oracle_run(cells, size, born, sustain, steps):
    field = O.Field(size, 26, np.uint8, cells)
    hashes = O.lifelike(field, O.moore_deltas(3), O.moore_rev(3), born,
                        sustain, steps, colors=False, hashes=True)
    return hashes, field


@pytest.mark.parametrize("size", [(32, 5, 7), (64, 17, 9), (128, 24, 40),
                                  (256, 9, 3), (2048, 3, 4), (32, 1, 1)])
@pytest.mark.parametrize("rule", [(B5, S45), (1 << 6, (1 << 5) | (1 << 7)),
                                  ((1 << 4) | (1 << 9), 0x3FF0)])
def test_small_vs_oracle(size, rule):
    rng = np.random.default_rng(sum(size))
    cells = (rng.random(int(np.prod(size))) < 0.2).astype(np.uint8)
    steps = 6
    want, field = oracle_run(cells, size, rule[0], rule[1], steps)
    bits = Life3DField(size, *rule)
    bits.set_cells(cells)
    u8 = Life3DBytes(size, *rule)
    u8.set_cells(cells)
    direct = Life3DBytes(size, *rule, direct=True)
    direct.set_cells(cells)
    for step in range(steps):
        bits.step()
        u8.step()
        direct.step()
        assert bits.crc32() == want[step], "bits step %d" % (step + 1)
        assert u8.crc32() == want[step], "u8 step %d" % (step + 1)
        assert direct.crc32() == want[step], "direct step %d" % (step + 1)


def test_dsl_model_matches_oracle():
    """The DSL 3-D model (plain engine) reproduces the oracle."""
    class Small3D(Life3DExperiment):
        size = (24, 20, 16)
        engine = "plain"
    model = LifeLike3D(Small3D)
    cells = model.gpu.arrays.cells.get()[:model.cells_num]
    want, field = oracle_run(cells, Small3D.size, B5, S45, 8)
    for step in range(8):
        model.step()
        got = model.gpu.arrays.cells.get()[:model.cells_num]
        assert O.crc32(got) == want[step]


def test_seed_chunks_equal_pattern():
    class Small3D(Life3DExperiment):
        size = (40, 30, 20)
    model = LifeLike3D(Small3D)
    cells = model.gpu.arrays.cells.get()[:model.cells_num]
    assert np.array_equal(cells, life3d_cells(Small3D.size, chunk=999))


def test_full_size_bits_vs_u8():
    """
    1024^3, 10 % density, 200 steps: the packed kernel, the uint8 plane
    stepped packed (packed once, expanded at every read) and the direct
    26-load stencil agree every 25 steps.
    """
    size = (1024, 1024, 1024)
    cells = life3d_cells(size)
    fields = [Life3DField(size, B5, S45), Life3DBytes(size, B5, S45),
              Life3DBytes(size, B5, S45, direct=True)]
    for field in fields:
        field.set_cells(cells)
    del cells
    for block in range(8):
        for field in fields:
            field.step(25)
        crcs = [field.crc32() for field in fields]
        if not crcs[0] == crcs[1] == crcs[2]:
            planes = [field.get_cells().reshape(size[::-1]) for field in fields]
            for a, b in ((0, 1), (0, 2), (1, 2)):
                diff = np.argwhere(planes[a] != planes[b])
                print("fields %d/%d differ in %d cells, first (z, y, x) %s"
                      % (a, b, len(diff), diff[:4].tolist()))
        assert crcs[0] == crcs[1] == crcs[2], "after %d steps: %s" % (
            25 * (block + 1), crcs)


@pytest.mark.parametrize("size", [(1024, 16, 128), (512, 24, 70)])
def test_wide_rows_and_z_tiles_vs_oracle(size):
    """Rows of 16-32 lane words and depth >= 64 (several z tiles)."""
    rng = np.random.default_rng(size[2])
    cells = (rng.random(int(np.prod(size))) < 0.15).astype(np.uint8)
    want, _ = oracle_run(cells, size, B5, S45, 4)
    fields = [Life3DField(size, B5, S45), Life3DBytes(size, B5, S45)]
    for field in fields:
        field.set_cells(cells)
    for step in range(4):
        for idx, field in enumerate(fields):
            field.step()
            assert field.crc32() == want[step], (idx, step + 1)


def test_dsl_1024_cube_on_the_bits_engine():
    """LifeLike3D at 1024^3 through ``model.step()``: the default engine is
    the bit-sliced one (only the main plane is on the device), and its
    state equals the packed kernel's run from the same cells every 50
    steps for 200 steps."""
    from xentica_amd.runtime.bits_engine import BitsEngine
    print("LifeLike3D 1024^3 through the DSL", flush=True)
    model = LifeLike3D(Life3DExperiment)
    assert isinstance(model._engine, BitsEngine)
    size = Life3DExperiment.size
    field = Life3DField(size, B5, S45)
    field.set_cells(life3d_cells(size))
    for block in range(4):
        for _ in range(50):
            model.step()
        field.step(50)
        assert model._engine.crc32() == field.crc32(), 50 * (block + 1)
    assert not model.gpu.arrays.cells._buffer  # buffer planes never allocated
