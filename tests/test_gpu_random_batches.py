"""
Differential tests of the fused kernel: seeded random batches of Life-like
automata (random birth/survival masks, odd sizes down to fields smaller
than a tile, torus or static border) at random tiles and depths, every
instance against the plain engine (the reference's structure) run alone.
"""
import binascii

import numpy as np
import pytest

from examples.workloads import LifeLikeRule
from xentica_amd import core, seeds
from xentica_amd.runtime.fused import Batch, TUNE_TILES

pytestmark = pytest.mark.gpu


class LifeLikeStatic(LifeLikeRule):
    """The same rule inside a static border of live cells."""

    class Topology(LifeLikeRule.Topology):
        """Static border."""

        border = core.StaticBorder(1)


def experiments(rng, count):
    exps = []
    for idx in range(count):
        width = int(rng.integers(1, 200))
        height = int(rng.integers(1, 200))
        level = int(rng.integers(50, 700))
        exps.append(type("Random%d" % idx, (core.Experiment, ), {
            "word": int(rng.integers(0, 2 ** 31)),
            "size": (width, height), "zoom": 1, "pos": [0, 0],
            "rule": int(rng.integers(0, 512)) | (int(rng.integers(0, 512)) << 9),
            "seed": seeds.patterns.PrimordialSoup(vals={
                "state": (seeds.random.RandInt(0, 1023) + level) >> 10}),
        }))
    return exps


def crc(model):
    return binascii.crc32(model.gpu.arrays.cells.get()[:model.cells_num])


@pytest.mark.parametrize("case", range(8))
def test_random_batch_against_plain(case):
    rng = np.random.default_rng(1000 + case)
    cls = LifeLikeStatic if case % 2 else LifeLikeRule
    tile = TUNE_TILES[int(rng.integers(0, len(TUNE_TILES)))]
    depth = int(rng.integers(1, 9))
    steps = int(rng.integers(1, 3 * depth + 2))
    exps = experiments(rng, 6)
    batch = Batch([cls(exp) for exp in exps], tile=tile, depth=depth)
    batch.run(steps, graph=bool(case % 3))
    batch.synchronize()
    batch.scatter()
    for exp, model in zip(exps, batch.models):
        plain = cls(type("Plain" + exp.__name__, (exp, ), {"engine": "plain"}))
        plain.rule = exp.rule
        for _ in range(steps):
            plain.step()
        what = (case, exp.__name__, exp.size, tile, depth, steps)
        assert crc(model) == crc(plain), what
        assert np.array_equal(model.gpu.arrays.colors.get(),
                              plain.gpu.arrays.colors.get()), what
