"""
The host C++ rendering of generated code (``runtime/hostc.py``).

* :class:`PlainHost` runs the reference's plain structure on the CPU; it
  is pinned by the reference's own golden CRCs (``tests/core/test_base.py``,
  ``tests/core/test_properties.py``), the same ones the GPU engines meet.
* :class:`LifeLikeProbe` recognises Life-like models exactly: the rules it
  reports are the ones the models were written with, and models that are
  not Life-like (a neighbour counted twice, a colour that depends on the
  neighbours, a birth count that kills) are refused.
"""
import binascii

import numpy as np
import pytest

from examples.game_of_life import (
    GameOfLife, GameOfLifeStatic, GameOfLifeColor, GameOfLife6D,
    GOLExperiment, GOLExperiment2, GOLExperimentColor, LifelikeCA,
    DiamoebaExperiment)
from examples.shifting_sands import ShiftingSands, ShiftingSandsExperiment
from examples.noisetv import NoiseTV, NoiseTVExperiment
from examples.workloads import (
    LifeLikeRule, LifeLike3D, Life3DExperiment, batch_experiments)
from xentica_amd import core
from xentica_amd.core import color_effects
from xentica_amd.core.exceptions import XenticaException
from xentica_amd.runtime.hostc import PlainHost, LifeLikeProbe
from xentica_amd.tools.rules import LifeLike


def host_run(model, steps, before_step=None):
    host = PlainHost(model)
    host.load(model.gpu.arrays.cells.get(), model.gpu.arrays.colors.get())
    for step in range(steps):
        if before_step:
            before_step(model, step)
        host.step()
    return host


@pytest.mark.parametrize("cls,exp,steps,want", [
    (GameOfLife, GOLExperiment, 1000, 2981695958),
    (GameOfLifeStatic, GOLExperiment, 1000, 1098273940),
    (GameOfLifeColor, GOLExperimentColor, 1000, 532957133),
    (ShiftingSands, ShiftingSandsExperiment, 1000, 1639048270),
    (NoiseTV, NoiseTVExperiment, 100, 2773894957),
])
def test_plain_host_golden(cls, exp, steps, want):
    host = host_run(cls(exp), steps)
    assert binascii.crc32(host.cells) == want


def test_plain_host_interactive_rule():
    """``test_base.py:218-225``: the rule switches to B3/S23 at t=23."""
    def switch(model, step):
        if step == 23:
            model.rule = LifeLike.golly2int("B3/S23")
    model = LifelikeCA(DiamoebaExperiment)
    host = host_run(model, 1000, switch)
    assert binascii.crc32(host.cells) == 4016906756


def test_plain_host_six_dimensions():
    class GOLExperiment6DLite(GOLExperiment2):
        size = (64, 36, 3, 3, 3, 3)
    host = host_run(GameOfLife6D(GOLExperiment6DLite), 1000)
    assert binascii.crc32(host.cells) == 2742543959


def test_probe_recognises_the_reference_models():
    probe = LifeLikeProbe(GameOfLife(GOLExperiment))
    assert probe.rule(probe.values()) == (
        8, 12, ((0, 0, 0), (255, 255, 255)))
    probe = LifeLikeProbe(LifelikeCA(DiamoebaExperiment))
    rule = LifeLike.golly2int("B35678/S5678")
    born, survive = rule & 0x1FF, rule >> 9
    assert probe.rule((rule, ))[:2] == (born, born | survive)
    assert probe.rule((6152, ))[:2] == (8, 12)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_probe_finds_each_batch_rule(seed):
    """Random birth/survival masks: the probe reports born and born |
    survival (the reference's ``next = born | sustained & state``)."""
    for exp in batch_experiments(count=6, seed=seed, side_min=40,
                                 side_max=80):
        probe = LifeLikeProbe(LifeLikeRule(exp))
        born, survive = exp.rule & 0x1FF, exp.rule >> 9
        assert probe.rule((exp.rule, ))[:2] == (born, born | survive)


def test_probe_3d_reference_offsets():
    """The 3-D Moore list keeps the centre and drops (1, 1, 1): the count
    of a live cell includes itself, so B5/S45 keeps cells with 4 or 5."""
    class Small(Life3DExperiment):
        size = (32, 32, 32)
    probe = LifeLikeProbe(LifeLike3D(Small))
    assert probe.rule((1 << 5, (1 << 4) | (1 << 5)))[:2] == (
        1 << 5, (1 << 4) | (1 << 5))
    # a birth count live cells cannot reach (26 needs a dead cell with 26
    # live cells of 25 others) is dropped
    assert probe.rule(((1 << 26) | 1, 1 << 26))[:2] == (1, 1 << 26)


class _Weighted(GameOfLife):
    """Counts neighbour 0 twice: not totalistic."""

    def absorb(self):
        count = core.IntegerVariable()
        count += self.neighbors[0].buffer.state
        for i in range(len(self.buffers)):
            count += self.neighbors[i].buffer.state
        is_born = (8 >> count) & 1
        is_sustain = (12 >> count) & 1
        self.main.state = is_born | is_sustain & self.main.state


class _NeighbourColour(GameOfLife):
    """Colours by the neighbour count: not a function of the state."""

    @color_effects.ColorEffect
    def color(self):
        val = self.neighbors[0].buffer.state * 100 + self.main.state * 100
        return (val, val, val)


class _KillOnBirth(GameOfLife):
    """``next = state ? S(count) : B(count)`` with B3 not in S2: a live
    cell with three neighbours dies, which ``born | sustain & state``
    cannot express."""

    def absorb(self):
        count = core.IntegerVariable()
        for i in range(len(self.buffers)):
            count += self.neighbors[i].buffer.state
        dead = (8 >> count) & 1 & (self.main.state ^ 1)
        alive = (4 >> count) & 1 & self.main.state
        self.main.state = dead | alive


@pytest.mark.parametrize("cls", [_Weighted, _NeighbourColour,
                                 _KillOnBirth])
def test_probe_refuses_models_that_are_not_life_like(cls):
    probe = LifeLikeProbe(cls(GOLExperiment))
    assert probe.rule(probe.values()) is None


def test_probe_refuses_by_structure():
    for cls, exp in ((GameOfLifeStatic, GOLExperiment),
                     (GameOfLifeColor, GOLExperimentColor),
                     (NoiseTV, NoiseTVExperiment)):
        with pytest.raises(XenticaException):
            LifeLikeProbe(cls(exp))


def test_probe_agrees_with_the_plain_host_step():
    """One plain host step of a random field equals the probe's rule
    applied by NumPy (a check of the probe that does not use it)."""
    exp = batch_experiments(count=1, seed=17, side_min=64, side_max=64)[0]
    model = LifeLikeRule(exp)
    probe = LifeLikeProbe(model)
    born, sustain, _ = probe.rule((exp.rule, ))
    host = host_run(model, 1)
    cells = model.gpu.arrays.cells.get()[:model.cells_num].reshape(
        exp.size[1], exp.size[0]).astype(np.int64)
    count = sum(np.roll(np.roll(cells, dy, 0), dx, 1)
                for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                if (dx, dy) != (0, 0))
    want = ((born >> count) & 1) | ((sustain >> count) & 1) & cells
    assert np.array_equal(host.cells.reshape(cells.shape), want)
