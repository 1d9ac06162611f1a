"""
EnergyLife (the EvoLife-style model, 64-bit words, and its 128-bit
variant) on the CPU: the oracle's pins in both layouts, the seed packing
and the generated code.

The model has no reference counterpart, so no reference vector exists for
it; the oracle is pinned by what the mathematics fixes:

* exact conservation of total energy (the flux is antisymmetric);
* the min/max principle (each flux is at most 1/8 of the difference);
* single-spike closed forms for one step;
* reduction to B3/S23 Life (uniform genome 6152, zero thresholds, no
  mutation, energies >= 1) against the NumPy roll-sum Life;
* the xorshift32 (13, 17, 5) period 2**32 - 1 (Marsaglia 2003) and one
  advance per step;
* newborns take a live neighbour's genome (no mutation);
* the 64-bit and 128-bit layouts give the same fields step for step.
"""
import numpy as np
import pytest

from oracle import oracle
from examples.energy_life import EnergyLife, EnergyLifeSmall, EnergyLifeWide

LAYOUTS = pytest.mark.parametrize("wide", [False, True],
                                  ids=["64bit", "128bit"])


def small_field(energy, genome, mode, rng, wide=False):
    """A 2-D oracle field from ``(w1, w0)`` field arrays."""
    energy = np.asarray(energy)
    size = (energy.shape[1], energy.shape[0])
    words = oracle.energy_words(
        np.ravel(energy), np.ravel(np.broadcast_to(genome, energy.shape)),
        np.ravel(np.broadcast_to(mode, energy.shape)),
        np.ravel(np.broadcast_to(rng, energy.shape)), wide)
    return oracle.Field(size, 8, oracle.WIDE if wide else np.uint64, words)


def fields_of(field):
    w0, w1 = field.size
    return [v.reshape(w1, w0) for v in oracle.energy_fields(field.cells)]


def numpy_xorshift(x):
    x = x.astype(np.uint32)
    x ^= x << np.uint32(13)
    x ^= x >> np.uint32(17)
    x ^= x << np.uint32(5)
    return x


@LAYOUTS
def test_seed_matches_model(wide):
    model = (EnergyLifeWide if wide else EnergyLife)(EnergyLifeSmall)
    cells = model.gpu.arrays.cells.get()[:model.cells_num]
    field = oracle.energy_life_field(EnergyLifeSmall.size, wide=wide)
    assert cells.dtype.itemsize == (16 if wide else 8)
    assert cells.tobytes() == field.cells.tobytes()


def test_wide_pack_state_is_modular_sum():
    """128-bit ``pack_state`` = sum of value << shift modulo 2**128."""
    model = EnergyLifeWide(EnergyLifeSmall)
    rng = np.random.default_rng(3)
    state = {
        "energy": rng.integers(-2 ** 40, 2 ** 40, 64),
        "genome": rng.integers(-2 ** 62, 2 ** 62, 64),
        "mode": rng.integers(-3, 3, 64),
        "rng": rng.integers(0, 2 ** 33, 64),
    }
    packed = model.pack_state(state)
    shifts = {"energy": 0, "genome": 16, "mode": 34, "rng": 35}
    for idx in range(64):
        want = sum(int(state[name][idx]) << shifts[name]
                   for name in state) % 2 ** 128
        got = int(packed['lo'][idx]) | (int(packed['hi'][idx]) << 64)
        assert got == want


def test_generated_code_uses_one_64_bit_word():
    model = EnergyLife(EnergyLifeSmall)
    src = model.cuda_source
    assert "unsigned long *fld" in src or "unsigned long long *fld" in src
    assert model.main.bit_width == 59
    assert "_cell_rng = (_cell_rng ^ (_cell_rng << 13));" in src
    assert "_cell_rng = (_cell_rng ^ (_cell_rng >> 17));" in src
    assert "_cell_rng = (_cell_rng ^ (_cell_rng << 5));" in src
    assert "<< 27)" in src


def test_generated_code_of_the_wide_variant_uses_one_128_bit_word():
    model = EnergyLifeWide(EnergyLifeSmall)
    src = model.cuda_source
    assert "unsigned __int128 *fld" in src
    assert model.main.bit_width == 67
    assert "<< 35)" in src


@LAYOUTS
def test_energy_conserved_and_bounded(wide):
    field = oracle.energy_life_field((96, 80), wide=wide)
    energy0 = fields_of(field)[0]
    total, low, high = energy0.sum(), energy0.min(), energy0.max()
    for _ in range(10):
        oracle.energy_life(field, 10)
        energy = fields_of(field)[0]
        assert energy.sum() == total
        assert low <= energy.min() and energy.max() <= high


@LAYOUTS
@pytest.mark.parametrize("centre_mode,shift", [(0, 3), (1, 4)])
def test_single_spike(centre_mode, shift, wide):
    """Spike E on an empty dead field with no births: one step."""
    big = 1000 if wide else 250
    energy = np.zeros((9, 11), dtype=np.int64)
    energy[4, 5] = big
    mode = np.zeros_like(energy)
    mode[4, 5] = centre_mode
    # genome 0: nothing is born; the spike (if live) dies at once
    field = small_field(energy, 0, mode, 1, wide)
    oracle.energy_life(field, 1, sustain=0, birth=0, mutation=0)
    new, genome, new_mode, _ = fields_of(field)
    share = big >> shift
    want = np.zeros_like(energy)
    want[3:6, 4:7] = share
    want[4, 5] = big - 8 * share
    assert np.array_equal(new, want)
    assert not new_mode.any() and not genome.any()


@LAYOUTS
def test_uniform_field_is_fixed(wide):
    energy = np.full((8, 8), 77, dtype=np.int64)
    field = small_field(energy, 0, 0, 5, wide)
    oracle.energy_life(field, 3)
    assert np.array_equal(fields_of(field)[0], energy)


@LAYOUTS
def test_reduces_to_life(wide):
    """Uniform B3/S23 genome, zero thresholds, no mutation: mode is Life."""
    rng = np.random.default_rng(7)
    energy = rng.integers(1, 256, (48, 64))
    mode = rng.integers(0, 2, (48, 64))
    field = small_field(energy, oracle.golly2int("B3/S23"), mode, 12345,
                        wide)
    state = mode.astype(np.uint8)
    for _ in range(30):
        oracle.energy_life(field, 1, sustain=0, birth=0, mutation=0)
        state = oracle.numpy_life_step(state)
        assert np.array_equal(fields_of(field)[2], state)


@LAYOUTS
def test_generator_advances_once_per_step(wide):
    field = oracle.energy_life_field((32, 24), wide=wide)
    rng0 = fields_of(field)[3]
    oracle.energy_life(field, 5)
    want = rng0
    for _ in range(5):
        want = numpy_xorshift(want)
    assert np.array_equal(fields_of(field)[3], want.astype(np.int64))


def test_layouts_agree():
    """The 64-bit and 128-bit layouts hold the same fields after every
    step of the seeded workload (energies stay within 0-255), colours
    included."""
    narrow = oracle.energy_life_field((64, 48))
    wide = oracle.energy_life_field((64, 48), wide=True)
    for _ in range(25):
        oracle.energy_life(narrow, 2)
        oracle.energy_life(wide, 2)
        for a, b in zip(fields_of(narrow), fields_of(wide)):
            assert np.array_equal(a, b)
        assert np.array_equal(narrow.col, wide.col)


def test_xorshift32_period():
    assert oracle.xorshift32_period(1) == 2 ** 32 - 1


def test_newborns_inherit_a_live_neighbours_genome():
    field = oracle.energy_life_field((40, 40))
    for _ in range(5):
        _, genome0, mode0, _ = fields_of(field)
        oracle.energy_life(field, 1, mutation=0)
        _, genome1, mode1, _ = fields_of(field)
        born = np.argwhere((mode0 == 0) & (mode1 == 1))
        assert len(born)
        for y, x in born:
            parents = {genome0[(y + dy) % 40, (x + dx) % 40]
                       for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                       if (dy or dx) and mode0[(y + dy) % 40, (x + dx) % 40]}
            assert genome1[y, x] in parents
