"""
EnergyLife: an EvoLife-style automaton whose total energy is conserved.

No counterpart in the reference; it is written with the reference's DSL
the way ``examples/evolife.py`` is (a Life-like rule carried per cell as
a genome, births that inherit a neighbour's genome, a per-cell RNG), and
adds what ``EvoLife`` does not have: exact energy conservation and a
32-bit xorshift generator.

State of a cell (59 bits, so the container is one 64-bit word):

* ``energy`` -- 8 bits, seeded uniform 0-255 (a cell's new energy lies
  between the smallest and the largest energy of its neighbourhood, see
  below, so energies never leave the seeded range);
* ``genome`` -- 18 bits, birth counts in bits 0-8 and survival counts in
  bits 9-17 (the encoding of ``LifeLike.golly2int``);
* ``mode`` -- 1 bit, the cell is alive;
* ``rng`` -- xorshift32 state, seeded per cell with a non-zero value.

One step:

* ``emit`` publishes energy, genome and mode in every buffer;
* ``absorb`` moves energy across each of the 8 edges by the flux
  ``sign(d) * (|d| >> s)``, ``d`` the energy difference and
  ``s = 3 + mode_x + mode_y`` (live cells hold their energy); both ends
  of an edge compute the same flux, so the total is conserved exactly,
  and since every flux is at most 1/8 of the difference a cell's new
  energy stays between the minimum and maximum of its neighbourhood
  (8 bits never overflow for energies seeded in 0-255);
* a live cell survives if its genome's survival bit for its live-
  neighbour count is set and its new energy is at least
  ``sustain_energy``; a dead cell is born if the genome of its richest
  live neighbour (first in neighbour order on ties) has the birth bit for
  the count and its new energy is at least ``birth_energy``; a newborn
  takes that genome, with one bit flipped when the cell's generator
  (advanced once per step) says so (probability ``mutation / 65536``).
"""
from xentica import core
from xentica import seeds
from xentica.core import color_effects

from examples.base import RegularCA


class EnergyLife(RegularCA):
    """Conserved energy, evolving Life-like genomes, xorshift32 per cell."""

    energy = core.IntegerProperty(max_val=2 ** 8 - 1)
    genome = core.IntegerProperty(max_val=2 ** 18 - 1)
    mode = core.IntegerProperty(max_val=1)
    rng = core.XorShiftProperty()
    sustain_energy = core.Parameter(default=32)
    birth_energy = core.Parameter(default=128)
    mutation = core.Parameter(default=16)

    def emit(self):
        """Every buffer carries energy, genome and mode."""
        for buf in self.buffers:
            buf.energy = self.main.energy
            buf.genome = self.main.genome
            buf.mode = self.main.mode
            buf.rng = 0

    def absorb(self):
        """Energy flux over every edge, then the cell's Life-like rule."""
        energy = core.IntegerVariable()
        energy += self.main.energy
        mode = core.IntegerVariable()
        mode += self.main.mode
        new_energy = core.IntegerVariable()
        new_energy += energy
        count = core.IntegerVariable()
        best_energy = core.IntegerVariable()
        best_energy += 0  # declared here: the loop reads it before writing
        best_genome = core.IntegerVariable()
        for i in range(len(self.buffers)):
            nbr = self.neighbors[i].buffer
            shift = 3 + mode + nbr.mode
            new_energy += (nbr.energy > energy) * (
                (nbr.energy - energy) >> shift)
            new_energy -= (energy > nbr.energy) * (
                (energy - nbr.energy) >> shift)
            count += nbr.mode
            # genome first: both updates test the previous best energy
            best_genome += nbr.mode * (nbr.energy > best_energy) * (
                nbr.genome - best_genome)
            best_energy += nbr.mode * (nbr.energy > best_energy) * (
                nbr.energy - best_energy)
        rnd = core.IntegerVariable()
        rnd += self.main.rng
        survive = ((self.main.genome >> 9 >> count) & 1) * (
            new_energy >= self.sustain_energy)
        born = ((best_genome >> count) & 1) * (best_energy > 0) * (
            new_energy >= self.birth_energy)
        new_mode = core.IntegerVariable()
        new_mode += mode * survive + (1 - mode) * born
        flip = ((rnd & 65535) < self.mutation) << ((rnd >> 16) % 18)
        newborn = (1 - mode) * new_mode
        genome = core.IntegerVariable()
        genome += self.main.genome
        genome += newborn * ((best_genome ^ flip) - genome)
        self.main.energy = new_energy
        self.main.mode = new_mode
        self.main.genome = genome

    @color_effects.MovingAverage
    def color(self):
        """Live cells red, energy green, survival genes blue."""
        red = self.main.mode * 255
        green = core.IntegerVariable()
        green += self.main.energy
        blue = (self.main.genome >> 10) & 255
        return (red, green, blue)


class EnergyLifeWide(EnergyLife):
    """The same model with a 16-bit energy field: 67 bits of state, so one
    128-bit word per cell (the engines' wide-word path; energies above
    255 are representable, the dynamics are unchanged)."""

    energy = core.IntegerProperty(max_val=2 ** 16 - 1)


class EnergyLifeExperiment(core.Experiment):
    """4096x4096 torus, word 42, uniform energies, genomes and modes."""

    word = 42
    size = (4096, 4096)
    zoom = 1
    pos = [0, 0]
    seed = seeds.patterns.PrimordialSoup(vals={
        "energy": seeds.random.RandInt(0, 255),
        "genome": seeds.random.RandInt(0, 2 ** 18 - 1),
        "mode": seeds.random.RandInt(0, 1),
        "rng": seeds.random.RandInt(1, 2 ** 32 - 1),
    })


class EnergyLifeSmall(EnergyLifeExperiment):
    """The same seeding on a 96x80 torus (parity tests)."""

    size = (96, 80)


def field_position(model, name):
    """``(shift, width)`` of property ``name`` in the main container (the
    fields in declaration order from bit 0)."""
    shift = 0
    for prop_name, prop in model.main.properties.items():
        if prop_name == name:
            return shift, prop.bit_width
        shift += prop.bit_width
    raise KeyError(name)


def field_values(model, name):
    """The values of property ``name`` of every cell of ``model``'s main
    plane (the container's bit layout: fields in declaration order from
    bit 0; 128-bit words as two little-endian uint64 halves)."""
    import numpy as np
    shift, width = field_position(model, name)
    words = model.gpu.arrays.cells.get()[:model.cells_num]
    if words.dtype.names is None:
        val = words.astype(np.uint64) >> np.uint64(shift)
        return (val & np.uint64((1 << width) - 1)).astype(np.int64)
    lo = words['lo'].astype(np.uint64)
    hi = words['hi'].astype(np.uint64)
    if shift >= 64:
        val = hi >> np.uint64(shift - 64)
    elif shift:
        val = (lo >> np.uint64(shift)) | (hi << np.uint64(64 - shift))
    else:
        val = lo
    return (val & np.uint64((1 << width) - 1)).astype(np.int64)
