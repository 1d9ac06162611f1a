"""Seeds (``tests/seeds/test_random.py``, ``test_patterns.py`` of the reference)."""
import binascii

import numpy as np
import pytest

from xentica_amd.seeds.random import RandInt, LocalRandom
from xentica_amd.seeds.patterns import BigBang, PrimordialSoup, ValDict
from examples.game_of_life import GameOfLife, GOLExperiment

TEST_SEED = "test"


def test_standard_random_seed():
    rng = LocalRandom(TEST_SEED)
    assert [rng.standard.randint(11, 23) for _ in range(10)] == \
        [22, 15, 21, 23, 14, 14, 11, 20, 17, 23]


def test_numpy_rand_seed():
    rng = LocalRandom(TEST_SEED)
    assert list(rng.numpy.randint(11, 23, (10, ))) == \
        [19, 16, 22, 19, 15, 22, 20, 20, 13, 15]


class RandomHolder:
    def __init__(self, seed=None):
        self.random = LocalRandom(seed)


def test_randint_descriptor_seed():
    holder = RandomHolder(TEST_SEED)
    holder.__class__.rand_val = RandInt(11, 23)
    assert [holder.rand_val for _ in range(10)] == \
        [22, 15, 21, 23, 14, 14, 11, 20, 17, 23]


def test_valdict():
    vals = ValDict({'a': 2, 's': RandInt(11, 23), 'd': 3.3})
    assert vals['a'] == 2 and vals['d'] == 3.3
    assert 11 <= vals['s'] <= 23
    assert list(vals.keys()) == ['a', 's', 'd']
    with pytest.raises(KeyError):
        vals['missing']
    with pytest.raises(NotImplementedError):
        vals['a'] = 3


def test_pattern_expressions():
    """Unary and reflected seed arithmetic (``expressions.py:123-204``)."""
    holder = RandomHolder(TEST_SEED)
    holder.size = (10,)
    expr = (RandInt(0, 9) + 7) >> 4
    vals = expr.__get__(holder, RandomHolder)
    assert set(np.unique(vals)) <= {0, 1}
    assert list((0 + RandInt(1, 1)).__get__(holder, RandomHolder)) == [1] * 10
    assert list((-RandInt(2, 2)).__get__(holder, RandomHolder)) == [-2] * 10


class TestExperiment(GOLExperiment):
    size = (100, 100, )
    seed = BigBang(pos=(32, 20), size=(10, 10),
                   vals={"state": RandInt(0, 1)})


def test_bigbang_area():
    bsca = GameOfLife(TestExperiment)
    cells = bsca.gpu.arrays.cells.get()[:bsca.cells_num]
    for i in np.where(cells == 1)[0]:
        x, y = i % 100, i // 100
        assert 32 <= x < 42 and 20 <= y < 30


def test_bigbang_random_area_and_clamp():
    bsca = GameOfLife(TestExperiment)
    for pattern in (BigBang(pos=(10, 10), vals={'state': RandInt(0, 1)}),
                    BigBang(size=(10, 10), vals={'state': RandInt(0, 1)}),
                    BigBang(vals={'state': RandInt(0, 1)})):
        cells = np.zeros((10000, ), dtype=np.int32)
        pattern.generate(cells, bsca)
        assert np.sum(cells) > 0
    pattern = BigBang(pos=(90, 90), size=(20, 20),
                      vals={'state': RandInt(0, 1)})
    pattern.generate(np.zeros((10000, ), dtype=np.int32), bsca)
    assert list(pattern.pos) == [80, 80]


def test_primordial_soup_crc():
    """``test_patterns.py:155-164``."""
    bsca = GameOfLife(TestExperiment)
    seed = PrimordialSoup(vals={'state': RandInt(0, 1)})
    cells = np.zeros((10000, ), dtype=np.int32)
    seed.random = LocalRandom("test")
    seed.generate(cells, bsca)
    assert binascii.crc32(cells[:10000]) == 1648433356


def test_model_seed_matches_oracle():
    """The model's step-0 cells equal the oracle's restated seeding."""
    from oracle import oracle as O
    model = GameOfLife(GOLExperiment)
    cells = model.gpu.arrays.cells.get()
    assert cells.size == model.cells_num * 9 + 1
    assert np.array_equal(cells[:model.cells_num],
                          O.gol_experiment_field().cells)


def test_primordial_soup_in_chunks_equals_one_draw(monkeypatch):
    """A board larger than the chunk is drawn in chunks with the values of
    one draw of the whole board (the legacy stream continues), for a plain
    RandInt and for an expression of per-cell draws; a constant draw in
    the expression keeps the one-draw path."""
    import numpy as np
    from xentica_amd import seeds
    from examples.workloads import LifeLike3D, Life3DExperiment

    class Soup(Life3DExperiment):
        size = (64, 20, 10)

    exprs = [seeds.random.RandInt(0, 1),
             (seeds.random.RandInt(0, 9) + 7) >> 4,
             seeds.random.RandInt(0, 1) * seeds.random.RandInt(0, 1, True)]
    for expr in exprs:
        Soup.seed = seeds.patterns.PrimordialSoup(vals={"state": expr})
        whole = LifeLike3D(Soup).gpu.arrays.cells.get()
        monkeypatch.setattr(seeds.patterns.PrimordialSoup, "CHUNK", 999)
        Soup.seed = seeds.patterns.PrimordialSoup(vals={"state": expr})
        chunked = LifeLike3D(Soup).gpu.arrays.cells.get()
        monkeypatch.undo()
        assert np.array_equal(whole, chunked)
