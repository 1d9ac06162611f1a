"""
The fused engine (emit+absorb in one generated kernel, buffers in LDS,
several steps per launch) against the reference's golden checksums, the
plain engine and the oracle; and batches of uneven automata with their own
rules against single runs.
"""
import binascii

import numpy as np
import pytest

from examples.game_of_life import (
    GameOfLife, GameOfLifeStatic, GameOfLifeColor, GOLExperiment,
    GOLExperiment2, GOLExperimentColor, LifelikeCA, DiamoebaExperiment)
from examples.shifting_sands import ShiftingSands, ShiftingSandsExperiment
from examples.noisetv import NoiseTV, NoiseTVExperiment
from examples.evolife import (
    EvoLife, CrossbreedingExperiment, CrossbreedingExperiment2)
from examples.workloads import (
    Life512Experiment, LifeLikeRule, batch_experiments)
from xentica_amd.core import color_effects
from xentica_amd.runtime import hip
from xentica_amd.runtime.fused import Batch, FusedEngine, TUNE_TILES

pytestmark = pytest.mark.gpu


class PlainColorLife(GameOfLife):
    """Life whose colour is stored as computed (no moving average)."""

    @color_effects.ColorEffect
    def color(self):
        """Live cells white, dead ones dark blue."""
        val = self.main.state * 255
        return (val, val, 40 + val)


def fused(exp, engine="fused"):
    """The experiment with a fused engine selected."""
    return type(exp.__name__ + "Fused", (exp,), {"engine": engine})


def run(model, steps):
    for _ in range(steps):
        model.step()
    return model


def crc(model):
    return binascii.crc32(model.gpu.arrays.cells.get()[:model.cells_num])


@pytest.mark.parametrize("cls,exp,steps,want", [
    (GameOfLife, GOLExperiment, 1000, 2981695958),
    (GameOfLifeStatic, GOLExperiment, 1000, 1098273940),
    (GameOfLifeColor, GOLExperimentColor, 1000, 532957133),
    (LifelikeCA, DiamoebaExperiment, 1000, 4016906756),
    (ShiftingSands, ShiftingSandsExperiment, 1000, 1639048270),
    (NoiseTV, NoiseTVExperiment, 100, 2773894957),
])
@pytest.mark.parametrize("engine", ["fused", "fused_unfolded",
                                    "fused_generic"])
def test_golden(cls, exp, steps, want, engine):
    model = cls(fused(exp, engine))
    assert isinstance(model._engine, FusedEngine)
    assert model._engine.program.folded == (engine != "fused_unfolded")
    assert model._engine.program.split == (engine == "fused")
    assert crc(run(model, steps)) == want


def test_evolife_golden():
    """``test_base.py:197-216`` on the fused engine (uint64 words)."""
    for name, value in (("death_speed", 15), ("max_genes", 9),
                        ("mutation_prob", 0.0)):
        EvoLife.__dict__[name].__set__(None, value)
    model = run(EvoLife(fused(CrossbreedingExperiment),
                        legacy_coloring=False), 100)
    assert crc(model) == 2061779088
    assert binascii.crc32(model.gpu.arrays.colors.get()) == 68940598
    model = run(EvoLife(fused(CrossbreedingExperiment2)), 100)
    assert crc(model) == 3381214984
    assert binascii.crc32(model.gpu.arrays.colors.get()) == 2253830262


def test_render_after_fused_steps():
    class RenderExperiment(fused(GOLExperiment)):
        zoom = 1
    model = GameOfLife(RenderExperiment)
    model.set_viewport(RenderExperiment.size)
    run(model, 1000)
    assert binascii.crc32(model.render() / 255) == 1955702083


def test_colors_match_plain_every_step():
    plain = GameOfLifeColor(GOLExperimentColor)
    fast = GameOfLifeColor(fused(GOLExperimentColor))
    for step in range(40):
        plain.step()
        fast.step()
        assert crc(fast) == crc(plain), step
        assert np.array_equal(fast.gpu.arrays.colors.get(),
                              plain.gpu.arrays.colors.get()), step


@pytest.mark.parametrize("engine", ["fused", "fused_generic"])
@pytest.mark.parametrize("outside", [False, True])
def test_loaded_colours_and_the_last_step_form(engine, outside):
    """
    Life's colour is 255 * state with fades of 255: from stored colours in
    [0, 255] every step's colour is its new value, so the engine launches
    the form that computes the colour on a launch's last step only.
    Loaded colours within [0, 255] keep that form; rows of colours far
    outside it (and single cells just outside) make the engine compute
    every step's colour.  Both match the plain engine after launches of 1
    to 8 steps.
    """
    plain = GameOfLife(GOLExperiment2)
    fast = GameOfLife(fused(GOLExperiment2, engine))
    program = fast._engine.program
    assert program.color_plan == ((0, 255), ) * 3
    cells = plain.gpu.arrays.cells.get()
    rng = np.random.default_rng(5)
    colors = rng.integers(0, 256, plain.cells_num * 3).astype(np.int32)
    if outside:
        width = plain.size[0]
        colors[3 * width * 3:40 * width * 3] = rng.integers(
            -2000, 2000, 37 * width * 3)
        far = 200 * width * 3
        colors[far:far + 4] = (256, -1, 255, 0)
        colors[far + 3 * width * 20 + 7] = 1 << 30
    for model in (plain, fast):
        model._engine.load_state(cells, colors)
    assert program.admits(fast.gpu.arrays.colors.bounds) == (not outside)
    for count in (1, 8, 5, 16, 3):
        run(plain, count)
        run(fast, count)
        assert crc(fast) == crc(plain), count
        assert np.array_equal(fast.gpu.arrays.colors.get(),
                              plain.gpu.arrays.colors.get()), count
        want = None if outside else program.color_values
        assert fast.gpu.arrays.colors.bounds == want


def test_batch_with_colours_outside_the_plan():
    """One instance's loaded colours lie outside [0, 255]: the batch runs
    the every-step form, and every instance matches its plain twin."""
    exps = batch_experiments(count=4, seed=3, side_min=30, side_max=200)
    models = [LifeLikeRule(exp) for exp in exps]
    twins = [LifeLikeRule(fused(exp, "plain")) for exp in exps]
    odd = models[2]
    colors = np.random.default_rng(1).integers(
        -999, 999, odd.cells_num * 3).astype(np.int32)
    for model in (odd, twins[2]):
        model._engine.load_state(model.gpu.arrays.cells.get(), colors)
    batch = Batch(models, tile=(64, 16), depth=6)
    assert not batch.program.admits(batch.color_bounds)
    batch.run(20)
    batch.synchronize()
    batch.scatter()
    for exp, model, twin in zip(exps, models, twins):
        twin.rule = exp.rule
        run(twin, 20)
        assert crc(model) == crc(twin), exp.__name__
        assert np.array_equal(model.gpu.arrays.colors.get(),
                              twin.gpu.arrays.colors.get()), exp.__name__
        assert model.gpu.arrays.colors.bounds is None


def test_set_after_deferred_steps():
    """Steps the fused engine has deferred run before a ``set()`` of the
    cells or colours, as the plain engine's steps have."""
    plain = GameOfLife(GOLExperiment2)
    fast = GameOfLife(fused(GOLExperiment2))
    rng = np.random.default_rng(11)
    cells = plain.gpu.arrays.cells.get()
    cells[:plain.cells_num] = rng.integers(0, 2, plain.cells_num)
    colors = rng.integers(0, 256, plain.cells_num * 3).astype(np.int32)
    for model in (plain, fast):
        run(model, 3)
        model.gpu.arrays.cells.set(cells)
        model.gpu.arrays.colors.set(colors)
        run(model, 5)
    assert crc(fast) == crc(plain)
    assert np.array_equal(fast.gpu.arrays.colors.get(),
                          plain.gpu.arrays.colors.get())


def test_memoryless_colour_only_last_step():
    """A plain colour effect never reads the stored colour: the fused
    kernel computes it on a launch's last step and does not load it."""
    plain = PlainColorLife(GOLExperiment2)
    fast = PlainColorLife(fused(GOLExperiment2))
    assert fast._engine.program.color_plan == ()
    assert "#define _CFAST 1" in fast._engine.program.source_fast
    for count in (1, 8, 13):
        run(plain, count)
        run(fast, count)
        assert crc(fast) == crc(plain), count
        assert np.array_equal(fast.gpu.arrays.colors.get(),
                              plain.gpu.arrays.colors.get()), count


def test_life512_deferred_vs_oracle():
    """1000 deferred steps, checked against the oracle every 100."""
    from oracle import oracle as O
    model = GameOfLife(fused(Life512Experiment))
    start = model.gpu.arrays.cells.get()[:model.cells_num]
    field = O.Field((512, 512), 8, np.uint8, start)
    for block in range(10):
        run(model, 100)
        O.lifelike(field, O.moore_deltas(2), O.moore_rev(2), 8, 12, 100,
                   colors=False)
        assert np.array_equal(model.gpu.arrays.cells.get()[:model.cells_num],
                              field.cells), block


def test_save_load_fused(tmp_path):
    path = str(tmp_path / "fused.ca")
    ca1 = run(GameOfLife(fused(GOLExperiment)), 500)
    ca1.save(path)
    ca2 = GameOfLife(fused(GOLExperiment))
    ca2.load(path)
    assert crc(run(ca2, 500)) == 2981695958


def test_batch_matches_single_runs():
    """Uneven sizes and per-instance rules: batch == each model alone."""
    exps = batch_experiments(count=12, seed=7, side_min=20, side_max=300)
    batch_models = [LifeLikeRule(exp) for exp in exps]
    batch = Batch(batch_models)
    batch.run(37)
    batch.synchronize()
    batch.scatter()
    for exp, model in zip(exps, batch_models):
        single = LifeLikeRule(exp)
        single.rule = exp.rule
        run(single, 37)
        assert crc(model) == crc(single), exp.__name__
        assert np.array_equal(model.gpu.arrays.colors.get(),
                              single.gpu.arrays.colors.get())


@pytest.mark.parametrize("tile", TUNE_TILES)
@pytest.mark.parametrize("depth", [1, 3, 8])
def test_batch_split_loop_every_tile(tile, depth):
    """The split step loop (column-owning threads, ring of halo cells)
    against the generic loop, for every tile the autotuner tries, with a
    step count that leaves a partial launch at the end."""
    exps = batch_experiments(count=5, seed=depth, side_min=9, side_max=150)
    fast = Batch([LifeLikeRule(exp) for exp in exps], tile=tile, depth=depth)
    slow = Batch([LifeLikeRule(exp) for exp in exps], tile=tile, depth=depth,
                 split=False)
    assert fast.program.split and not slow.program.split
    for batch in (fast, slow):
        batch.run(2 * depth + 1, graph=False)
        batch.synchronize()
        batch.scatter()
    for a, b in zip(fast.models, slow.models):
        assert crc(a) == crc(b)
        assert np.array_equal(a.gpu.arrays.colors.get(),
                              b.gpu.arrays.colors.get())


@pytest.mark.parametrize("tile", TUNE_TILES)
def test_static_border_split_loop_every_tile(tile):
    """Static border (live cells outside the field) through the split step
    loop at every autotuner tile: off-board cells of edge tiles, inner and
    ring, against the plain engine run alone."""
    from xentica_amd import seeds
    sizes = [(70, 45), (33, 130), (9, 11)]
    exps = [type("Static%d" % idx, (GOLExperiment, ), {
        "size": size, "word": 5 + idx,
        "seed": seeds.patterns.PrimordialSoup(
            vals={"state": seeds.random.RandInt(0, 1)})})
        for idx, size in enumerate(sizes)]
    for depth in (3, 8):
        batch = Batch([GameOfLifeStatic(exp) for exp in exps], tile=tile,
                      depth=depth)
        assert batch.program.split and batch.program.static
        batch.run(2 * depth + 1, graph=False)
        batch.synchronize()
        batch.scatter()
        for exp, model in zip(exps, batch.models):
            plain = run(GameOfLifeStatic(
                type("Plain" + exp.__name__, (exp, ), {"engine": "plain"})),
                2 * depth + 1)
            assert crc(model) == crc(plain), (exp.__name__, depth)
            assert np.array_equal(model.gpu.arrays.colors.get(),
                                  plain.gpu.arrays.colors.get())


def test_batch_rule_change_between_graph_replays():
    """Per-instance parameter values changed between runs of a captured
    graph take effect (the table is re-uploaded, the graph reads it)."""
    exps = batch_experiments(count=4, seed=9, side_min=30, side_max=120)
    models = [LifeLikeRule(exp) for exp in exps]
    batch = Batch(models, tile=(64, 16), depth=3)
    plain = [LifeLikeRule(type("P" + exp.__name__, (exp, ), {"engine": "plain"}))
             for exp in exps]
    rules = [exp.rule for exp in exps]
    schedule = [None, 6152, None, 0b111000 | (0b1100 << 9)]
    for new_rule in schedule:
        if new_rule is not None:
            models[1].rule = new_rule
            rules[1] = new_rule
        batch.run(7, graph=True)
        # a parameter is class-wide, as in the reference: a single model
        # steps with the value last set on the class
        for model, rule in zip(plain, rules):
            model.rule = rule
            for _ in range(7):
                model.step()
    batch.synchronize()
    batch.scatter()
    for a, b in zip(models, plain):
        assert crc(a) == crc(b)


def test_batch_graph_replays():
    """The captured graph replayed several times equals eager launches."""
    exps = batch_experiments(count=6, seed=3, side_min=40, side_max=200)
    eager = Batch([LifeLikeRule(exp) for exp in exps])
    graphed = Batch([LifeLikeRule(exp) for exp in exps])
    for _ in range(3):
        eager.run(20, graph=False)
        graphed.run(20, graph=True)
        eager.synchronize()
        graphed.synchronize()
        eager.scatter()
        graphed.scatter()
        for a, b in zip(eager.models, graphed.models):
            assert crc(a) == crc(b)


def test_batch_full_workload_vs_plain_and_oracle():
    """The 256-instance batch, 1000 steps, at the launch the autotuner picks
    for it on MI355X (64x32 tiles, depth 8, graph): every instance equal to
    the plain
    engine run alone with its own rule, and the eight smallest also equal
    to the oracle (the reference's plain structure on the CPU) with their
    own birth/survival masks."""
    from oracle import oracle as O
    from xentica_amd.runtime import kernels
    exps = batch_experiments()
    models = [LifeLikeRule(exp) for exp in exps]
    starts = [m.gpu.arrays.cells.get()[:m.cells_num] for m in models]
    batch = Batch(models, tile=(64, 32), depth=8)
    print("batch of %d, 1000 steps" % len(models), flush=True)
    batch.run(1000)
    batch.synchronize()
    batch.scatter()
    got = [kernels.crc32_u8(m.gpu.arrays.cells.device_ptr(), m.cells_num)
           for m in models]
    singles = [LifeLikeRule(type("Plain%d" % idx, (exp,), {"engine": "plain"}))
               for idx, exp in enumerate(exps)]
    # one plain module per field size: compiled in worker processes first
    hip.precompile([m.cuda_source for m in singles],
                   hip.compile_options(LifeLikeRule.fp_contract))
    for idx, (exp, single) in enumerate(zip(exps, singles)):
        single.rule = exp.rule
        run(single, 1000)
        want = kernels.crc32_u8(single.gpu.arrays.cells.device_ptr(),
                                single.cells_num)
        assert got[idx] == want, "instance %d vs plain" % idx
    sizes = [exp.size[0] for exp in exps]
    for idx in np.argsort(sizes, kind="stable")[:8]:
        exp = exps[idx]
        field = O.Field(exp.size, 8, np.uint8, starts[idx])
        O.lifelike(field, O.moore_deltas(2), O.moore_rev(2),
                   exp.rule & 0x1FF, (exp.rule >> 9) & 0x1FF, 1000,
                   colors=False)
        assert got[idx] == binascii.crc32(field.cells), \
            "instance %d vs oracle" % idx


def test_batch_parameter_change_while_a_replay_runs():
    """A rule changed between two back-to-back runs, with no host sync in
    between, on instances large enough that the first replay is still
    running when the table is uploaded: the second run sees the new rule
    and the first the old one (the upload is ordered on the batch's
    stream)."""
    exps = batch_experiments(count=3, seed=5, side_min=1900, side_max=2048)
    models = [LifeLikeRule(exp) for exp in exps]
    batch = Batch(models, tile=(64, 16), depth=8)
    rules = [exp.rule for exp in exps]
    new_rule = 6152 if rules[0] != 6152 else 0b111000 | (0b1100 << 9)
    batch.run(96)
    models[0].rule = new_rule
    batch.run(96)
    batch.synchronize()
    batch.scatter()
    for idx, (model, exp) in enumerate(zip(models, exps)):
        plain = LifeLikeRule(type("P%d" % idx, (exp, ), {"engine": "plain"}))
        plain.rule = exp.rule
        run(plain, 96)
        plain.rule = new_rule if idx == 0 else exp.rule
        run(plain, 96)
        assert crc(model) == crc(plain), idx


@pytest.mark.parametrize("size", [(5, 3), (1, 7), (33, 2), (70, 65)])
def test_fields_smaller_than_a_tile(size):
    """Fields smaller than the tile and its halo wrap several times."""
    from xentica_amd import seeds

    def experiment(engine):
        return type("Tiny%s" % engine, (core_experiment(), ), {
            "size": size, "engine": engine, "word": sum(size),
            "seed": seeds.patterns.PrimordialSoup(
                vals={"state": seeds.random.RandInt(0, 1)})})

    fused_model = GameOfLife(experiment("fused"))
    plain_model = GameOfLife(experiment("plain"))
    for step in range(9):
        fused_model.step()
        plain_model.step()
        assert crc(fused_model) == crc(plain_model), step
    assert np.array_equal(fused_model.gpu.arrays.colors.get(),
                          plain_model.gpu.arrays.colors.get())


def core_experiment():
    from xentica_amd import core
    return core.Experiment
