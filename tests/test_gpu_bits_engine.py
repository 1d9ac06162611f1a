"""
The ``bits`` engine (Life-like DSL models on the bit-sliced kernels)
through the reference's API: golden CRCs of ``tests/core/test_base.py``,
per-step hashes against the oracle, colours and renders against the plain
engine, the fall back to the fused engine for colours the proof does not
cover, 3-D models, and the headline workload through ``model.step()``.
"""
import binascii

import numpy as np
import pytest

from examples.game_of_life import (
    GameOfLife, GOLExperiment, GOLExperiment2, LifelikeCA,
    DiamoebaExperiment)
from examples.workloads import (
    Life512Experiment, Life16kExperiment, LifeLike3D, Life3DExperiment,
    LifeLikeRule, batch_experiments)
from xentica_amd import core, seeds
from xentica_amd.runtime import kernels
from xentica_amd.runtime.bits_engine import BitsEngine

pytestmark = pytest.mark.gpu


def with_engine(exp, engine):
    return type(exp.__name__ + engine.title(), (exp, ), {"engine": engine})


def run(model, steps):
    for _ in range(steps):
        model.step()
    return model


def crc(model):
    return binascii.crc32(model.gpu.arrays.cells.get()[:model.cells_num])


def test_auto_selects_bits():
    model = GameOfLife(GOLExperiment)
    assert isinstance(model._engine, BitsEngine)


@pytest.mark.parametrize("cls,exp,want", [
    (GameOfLife, GOLExperiment, 2981695958),
    (LifelikeCA, DiamoebaExperiment, 4016906756),
])
def test_golden(cls, exp, want):
    model = cls(with_engine(exp, "bits"))
    assert isinstance(model._engine, BitsEngine)
    assert crc(run(model, 1000)) == want


def test_render_golden():
    class RenderExperiment(with_engine(GOLExperiment, "bits")):
        zoom = 1
    model = GameOfLife(RenderExperiment)
    model.set_viewport(RenderExperiment.size)
    run(model, 1000)
    assert binascii.crc32(model.render() / 255) == 1955702083


def test_two_models_interleaved_and_pause():
    one = GameOfLife(with_engine(GOLExperiment, "bits"))
    two = GameOfLife(with_engine(GOLExperiment, "bits"))
    for _ in range(500):
        one.step()
        two.step()
    two.paused = True
    for _ in range(500):
        one.step()
        two.step()
    two.paused = False
    run(two, 500)
    assert crc(one) == crc(two) == 2981695958


def test_save_load(tmp_path):
    path = str(tmp_path / "bits.ca")
    first = run(GameOfLife(with_engine(GOLExperiment, "bits")), 500)
    first.save(path)
    second = GameOfLife(with_engine(GOLExperiment, "bits"))
    second.load(path)
    assert crc(run(second, 500)) == 2981695958


def test_life512_every_step_vs_oracle():
    """The C1 workload: the hash of every one of 1000 steps (each read
    expands the packed state; the next step packs it again)."""
    from oracle import oracle as O
    model = GameOfLife(Life512Experiment)
    assert isinstance(model._engine, BitsEngine)
    start = model.gpu.arrays.cells.get()[:model.cells_num]
    field = O.Field((512, 512), 8, np.uint8, start)
    want = O.lifelike(field, O.moore_deltas(2), O.moore_rev(2), 8, 12, 1000,
                      colors=False, hashes=True)
    for step in range(1000):
        model.step()
        assert crc(model) == want[step], "step %d" % (step + 1)


def test_colours_and_cells_match_plain():
    plain = GameOfLife(with_engine(GOLExperiment2, "plain"))
    bits = GameOfLife(with_engine(GOLExperiment2, "bits"))
    for count in (1, 7, 8, 9, 20):
        run(plain, count)
        run(bits, count)
        assert crc(bits) == crc(plain), count
        assert np.array_equal(bits.gpu.arrays.colors.get(),
                              plain.gpu.arrays.colors.get()), count


def test_colours_outside_the_proof_fall_back_to_fused():
    """Stored colours outside [0, 255]: the steps run on the fused engine
    (every step's colour) until the colours are admitted again; cells and
    colours equal the plain engine's throughout."""
    plain = GameOfLife(with_engine(GOLExperiment2, "plain"))
    bits = GameOfLife(with_engine(GOLExperiment2, "bits"))
    rng = np.random.default_rng(3)
    colors = rng.integers(-900, 900, plain.cells_num * 3).astype(np.int32)
    cells = plain.gpu.arrays.cells.get()
    for model in (plain, bits):
        model._engine.load_state(cells, colors)
    for count in (1, 5, 12):
        run(plain, count)
        run(bits, count)
        assert crc(bits) == crc(plain), count
        assert np.array_equal(bits.gpu.arrays.colors.get(),
                              plain.gpu.arrays.colors.get()), count
    assert bits._engine._fallback is not None


def test_set_after_deferred_steps():
    plain = GameOfLife(with_engine(GOLExperiment2, "plain"))
    bits = GameOfLife(with_engine(GOLExperiment2, "bits"))
    rng = np.random.default_rng(11)
    cells = plain.gpu.arrays.cells.get()
    cells[:plain.cells_num] = rng.integers(0, 2, plain.cells_num)
    colors = rng.integers(0, 256, plain.cells_num * 3).astype(np.int32)
    for model in (plain, bits):
        run(model, 3)
        model.gpu.arrays.cells.set(cells)
        model.gpu.arrays.colors.set(colors)
        run(model, 13)
    assert crc(bits) == crc(plain)
    assert np.array_equal(bits.gpu.arrays.colors.get(),
                          plain.gpu.arrays.colors.get())


def test_rule_changes_between_steps():
    """An interactive rule changed every few steps (each value probed)."""
    exp = batch_experiments(count=1, seed=23, side_min=256,
                            side_max=256)[0]
    plain = LifeLikeRule(with_engine(exp, "plain"))
    bits = LifeLikeRule(with_engine(exp, "bits"))
    rng = np.random.default_rng(5)
    for _ in range(6):
        rule = int(rng.integers(0, 1 << 18))
        plain.rule = rule
        run(plain, 5)
        bits.rule = rule
        run(bits, 5)
        assert crc(bits) == crc(plain)


@pytest.mark.parametrize("size", [(64, 64, 64), (128, 40, 24)])
def test_3d_vs_plain(size):
    class Small(Life3DExperiment):
        pass
    Small.size = size
    plain = LifeLike3D(with_engine(Small, "plain"))
    bits = LifeLike3D(with_engine(Small, "bits"))
    assert isinstance(bits._engine, BitsEngine)
    for count in (1, 3, 6):
        run(plain, count)
        run(bits, count)
        assert crc(bits) == crc(plain), count
        assert np.array_equal(bits.gpu.arrays.colors.get(),
                              plain.gpu.arrays.colors.get())


def test_headline_through_model_step():
    """Life 16384^2 (word 42, 50 %) through ``GameOfLife.step()`` on the
    default engine -- the path the benchmark times -- against the plain
    engine every 250 steps for 2000 steps, by the device CRC."""
    model = GameOfLife(Life16kExperiment)
    assert isinstance(model._engine, BitsEngine)
    plain = GameOfLife(with_engine(Life16kExperiment, "plain"))
    print("Life 16384^2 through model.step(), 2000 steps", flush=True)
    for block in range(8):
        run(model, 250)
        run(plain, 250)
        want = kernels.crc32_u8(plain.gpu.arrays.cells.device_ptr(),
                                plain.cells_num)
        assert model._engine.crc32() == want, "after %d" % (250 * block + 250)


def test_headline_every_step_validation_run():
    """The headline workload's validation run: Life 16384^2 through
    ``GameOfLife.step()`` on the bits engine and on the plain engine, the
    CRC of the canonical uint8 plane compared after every one of 2000
    steps (each read launches the one step queued; the timed run's 8-step
    launches are checked by the test above)."""
    model = GameOfLife(Life16kExperiment)
    assert isinstance(model._engine, BitsEngine)
    plain = GameOfLife(with_engine(Life16kExperiment, "plain"))
    print("Life 16384^2, CRC after each of 2000 steps", flush=True)
    for step in range(2000):
        model.step()
        plain.step()
        want = kernels.crc32_u8(plain.gpu.arrays.cells.device_ptr(),
                                plain.cells_num)
        assert model._engine.crc32() == want, "step %d" % (step + 1)


@pytest.mark.parametrize("zoom,pos,viewport", [(1, [0, 0], (640, 360)),
                                               (3, [17, 5], (200, 120)),
                                               (2, [630, 350], (333, 77))])
def test_render_from_packed_state_vs_plain(zoom, pos, viewport):
    """``render()`` on the bits engine projects the packed state directly
    (no colour expansion) and gives the plain engine's image, for zooms,
    pans across the torus edge and viewports smaller than the field."""
    exps = [type("Render" + name, (with_engine(GOLExperiment2, name), ),
                 {"zoom": zoom, "pos": list(pos)}) for name in ("plain",
                                                                "bits")]
    plain, bits = GameOfLife(exps[0]), GameOfLife(exps[1])
    for model in (plain, bits):
        model.set_viewport(viewport)
        run(model, 21)
    assert bits._engine._stale  # the packed state is the current one
    assert np.array_equal(bits.render(), plain.render())
    assert bits._engine._stale  # rendering did not expand it


def test_render_3d_projection_vs_plain():
    class Small(Life3DExperiment):
        size = (64, 48, 20)
        zoom = 1
        pos = [0, 0, 0]
    plain = LifeLike3D(with_engine(Small, "plain"))
    bits = LifeLike3D(with_engine(Small, "bits"))
    for model in (plain, bits):
        model.set_viewport((64, 48))
        run(model, 5)
    assert np.array_equal(bits.render(), plain.render())


def soup(size, word=5):
    """A 50 % soup torus of ``size`` (any width)."""
    return type("Soup%dx%d" % tuple(size), (core.Experiment, ), {
        "word": word, "size": tuple(size), "zoom": 1, "pos": [0, 0],
        "seed": seeds.patterns.PrimordialSoup(
            vals={"state": seeds.random.RandInt(0, 1)})})


RAGGED = [(100, 60), (33, 40), (97, 131), (1000, 700)]


@pytest.mark.parametrize("size", RAGGED, ids=["%dx%d" % s for s in RAGGED])
def test_ragged_widths_vs_plain(size):
    """Widths that are not a multiple of 32 run on padded rows (the batch
    kernel as a batch of one): cells and colours equal the plain
    engine's after every group of steps."""
    plain = GameOfLife(with_engine(soup(size), "plain"))
    bits = GameOfLife(with_engine(soup(size), "bits"))
    assert bits._engine.rows is not None
    for count in (1, 7, 8, 9, 20, 3):
        run(plain, count)
        run(bits, count)
        assert crc(bits) == crc(plain), count
    assert np.array_equal(bits.gpu.arrays.colors.get(),
                          plain.gpu.arrays.colors.get())


def test_ragged_every_step_vs_oracle():
    from oracle import oracle as O
    model = GameOfLife(with_engine(soup((97, 131), 11), "bits"))
    start = model.gpu.arrays.cells.get()[:model.cells_num]
    field = O.Field((97, 131), 8, np.uint8, start)
    want = O.lifelike(field, O.moore_deltas(2), O.moore_rev(2), 8, 12, 60,
                      colors=False, hashes=True)
    for step in range(60):
        model.step()
        assert crc(model) == want[step], "step %d" % (step + 1)


def test_ragged_rule_changes_and_render():
    exp = [e for e in batch_experiments(count=40, seed=23, side_min=64,
                                        side_max=300)
           if e.size[0] % 32][0]
    plain = LifeLikeRule(with_engine(exp, "plain"))
    bits = LifeLikeRule(with_engine(exp, "bits"))
    assert bits._engine.rows is not None
    rng = np.random.default_rng(9)
    for _ in range(5):
        rule = int(rng.integers(0, 1 << 18))
        plain.rule = rule
        run(plain, 6)
        bits.rule = rule
        run(bits, 6)
        assert crc(bits) == crc(plain)
    for model in (plain, bits):
        model.set_viewport((150, 90))
    assert np.array_equal(bits.render(), plain.render())


def test_ragged_full_size():
    """A 4010x3000 torus (126 words per row, 22 cells of the last word
    unused) against the plain engine over 96 steps."""
    plain = GameOfLife(with_engine(soup((4010, 3000)), "plain"))
    bits = GameOfLife(with_engine(soup((4010, 3000)), "bits"))
    assert bits._engine.rows is not None
    run(plain, 96)
    run(bits, 96)
    assert bits._engine.crc32() == kernels.crc32_u8(
        plain.gpu.arrays.cells.device_ptr(), plain.cells_num)
