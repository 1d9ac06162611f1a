"""
Emit folding of the fused kernels, decided at translation (CPU only).

Buffer slot ``p`` folds to ``(main & keep) | const`` when emit stores only
copies of the cell's own main fields and integer constants; anything else
keeps the buffer planes.
"""
import pytest

from xentica_amd import core
from xentica_amd.runtime.fused import FusedProgram
from examples.base import RegularCA
from examples.game_of_life import GameOfLife, GOLExperiment
from examples.shifting_sands import ShiftingSands, ShiftingSandsExperiment
from examples.energy_life import EnergyLife, EnergyLifeSmall, EnergyLifeWide


class Flipped(RegularCA):
    """Emit stores a function of the state: not foldable."""

    state = core.IntegerProperty(max_val=1)

    def emit(self):
        for buf in self.buffers:
            buf.state = self.main.state ^ 1

    def absorb(self):
        count = core.IntegerVariable()
        for i in range(len(self.buffers)):
            count += self.neighbors[i].buffer.state
        self.main.state = count & 1

    def color(self):
        return (0, 0, 0)


class StaticTwo(GameOfLife):
    """A static border whose word has a bit emit does not copy."""

    marker = core.IntegerProperty(max_val=1)

    class Topology(GameOfLife.Topology):
        border = core.StaticBorder(3)

    def emit(self):
        for buf in self.buffers:
            buf.state = self.main.state


def _fold(cls, exp, **kwargs):
    return cls(exp).translate_tile(**kwargs)[2].fold


def test_life_folds_to_copies():
    fold = _fold(GameOfLife, GOLExperiment)
    assert fold == {slot: (1, 0) for slot in range(1, 9)}


def test_shifting_sands_folds_to_one_copy_and_zeros():
    fold = _fold(ShiftingSands, ShiftingSandsExperiment)
    # buffers[0] goes to slot rev[0] + 1 = 8; the others carry 0
    assert fold[8] == (1, 0)
    assert all(fold[slot] == (0, 0) for slot in range(1, 8))


@pytest.mark.parametrize("cls,kept", [(EnergyLife, 27), (EnergyLifeWide, 35)])
def test_wide_word_fold_keeps_energy_genome_mode(cls, kept):
    """64- and 128-bit words: buffers are the main word's energy, genome
    and mode bits (the generator field zeroed)."""
    fold = _fold(cls, EnergyLifeSmall)
    assert set(fold.values()) == {((1 << kept) - 1, 0)}
    program = FusedProgram(cls(EnergyLifeSmall))
    assert program.folded and program.halo_per_step == 1
    assert "_tb" not in program.source


def test_function_of_state_does_not_fold():
    assert _fold(Flipped, GOLExperiment) is None
    program = FusedProgram(Flipped(GOLExperiment))
    assert not program.folded and "__shared__ _word _tb" in program.source


def test_static_border_word_must_survive_the_fold():
    assert _fold(StaticTwo, GOLExperiment) is None


def test_fold_can_be_switched_off():
    assert _fold(GameOfLife, GOLExperiment, fold=False) is None
    program = FusedProgram(GameOfLife(GOLExperiment), fold=False)
    assert not program.folded


def test_main_copy_dropped_when_absorb_stores_main():
    program = FusedProgram(GameOfLife(GOLExperiment))
    assert not program.main_copy
    assert "_mp[_cur ^ 1][_li] = _mp[_cur][_li];" not in program.source


def test_color_guard_only_for_pure_colors():
    from examples.noisetv import NoiseTV, NoiseTVExperiment
    life = FusedProgram(GameOfLife(GOLExperiment))
    assert life.color_guard and "if (_inner) {" in life.source
    # the last-step form holds each step with and without the colour and
    # never reads the stored colours or the clamp's old colour
    assert "#define _CFAST 1" in life.source_fast
    assert "if (_last) {" in life.source_fast
    assert "old_col" not in life.source_fast and "old_col" in life.source
    # NoiseTV's colour advances the generator: it must run on every cell
    # and every step
    noise = FusedProgram(NoiseTV(NoiseTVExperiment))
    assert not noise.color_guard and "if (_inner" not in noise.source
    assert noise.color_plan is None and noise.source_fast is None


def test_direct_colour_store_only_where_it_raises_occupancy():
    from examples.workloads import LifeLikeRule, batch_experiments
    model = LifeLikeRule(batch_experiments(count=1)[0])
    # 64x32: 27 KB of colour staging holds a CU to four workgroups; without
    # it the planes need 7.7 KB and a CU holds eight
    wide = FusedProgram(model, tile=(64, 32), depth=8)
    assert wide.direct_colors
    assert wide.lds_bytes_fast == 2 * 80 * 48 < wide.lds_bytes
    assert "#define _CDIRECT 1" in wide.source_fast
    assert "_cg[_index(_gyo, _gxo)] = make_int3(new_r, new_g, new_b);" \
        in wide.source_fast
    # the every-step form always stages the colours
    assert "#define _CDIRECT 0" in wide.source
    # 64x16 already fits eight workgroups with the staging
    narrow = FusedProgram(model, tile=(64, 16), depth=8)
    assert not narrow.direct_colors
    assert narrow.lds_bytes_fast == narrow.lds_bytes
    assert "#define _CDIRECT 0" in narrow.source_fast
