"""Which engine ``auto`` picks (no GPU needed: selection only translates
and probes the generated code on the host)."""
import pytest

from examples.game_of_life import (
    GameOfLife, GameOfLifeStatic, GameOfLifeColor, GOLExperiment,
    GOLExperimentColor, LifelikeCA, DiamoebaExperiment)
from examples.noisetv import NoiseTV, NoiseTVExperiment
from examples.workloads import LifeLike3D, Life3DExperiment, LifeLikeRule
from xentica_amd.core.exceptions import XenticaException


class _Small3D(Life3DExperiment):
    size = (64, 32, 16)


class _Ragged(GOLExperiment):
    size = (624, 360)


@pytest.mark.parametrize("cls,exp,engine", [
    (GameOfLife, GOLExperiment, "bits"),
    (LifelikeCA, DiamoebaExperiment, "bits"),
    (LifeLike3D, _Small3D, "bits"),
    (GameOfLife, _Ragged, "bits"),
    (GameOfLifeStatic, GOLExperiment, "fused"),
    (GameOfLifeColor, GOLExperimentColor, "fused"),
    (NoiseTV, NoiseTVExperiment, "fused"),
])
def test_auto(cls, exp, engine):
    assert cls(exp)._engine.name == engine


def test_bits_refuses_what_it_cannot_run():
    class StaticBits(GOLExperiment):
        engine = "bits"
    with pytest.raises(XenticaException):
        GameOfLifeStatic(StaticBits)
    from xentica_amd import core, seeds

    class NarrowBits(core.Experiment):
        """Narrower than one 32-cell word."""
        word = 7
        size = (20, 40)
        engine = "bits"
        seed = seeds.patterns.PrimordialSoup(
            vals={"state": seeds.random.RandInt(0, 1)})
    with pytest.raises(XenticaException):
        GameOfLife(NarrowBits)
    assert GameOfLife(type("NarrowAuto", (NarrowBits, ),
                           {"engine": "auto"}))._engine.name == "fused"
