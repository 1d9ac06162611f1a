"""
ConservedLife: energy-conserving Life-like dynamics with evolving rules.

The model of the reference's ``examples/conserved_life.py`` (its rules are
those of ``ConservedLife.emit``/``absorb``, lines 65-146), restated with
this package's DSL so that a reference user finds it here.  The reference
holds no golden hash for it; what it checks is that the total energy
never changes (``test_energy_conservation``, lines 237-256), which the
GPU tests check here too.

* A cell is *full* when its energy reaches ``full_treshold``.
* Emit: a full cell with energy above the threshold, or whose own rule
  does not sustain its number of full neighbours, spreads energy; so does
  an empty cell that its rule does not let be born.  It passes
  ``min(energy, death_speed)`` to one neighbour: the first *birthing*
  neighbour, else the first empty one, scanning from a random start gate;
  the neighbour receives the passer's rule with the energy.
* Absorb: energy = kept energy + everything received; an empty cell takes
  a crossover of the rules it received; ``birthing`` = its rule's birth
  bit for the number of full neighbours, for empty cells.

Emit stores the main state (``new_energy`` and the generator), so the
model runs on the plain engine (the fused engine refuses such models).
"""
import random

from xentica import core
from xentica import seeds
from xentica.core import color_effects
from xentica.seeds.random import RandInt
from xentica.tools import xmath
from xentica.tools.color import GenomeColor
from xentica.tools.genetics import genome_crossover
from xentica.tools.rules import LifeLike

from examples.base import RegularCA, RegularExperiment

ENERGY_MASK = 2 ** 14 - 1


class ConservedLife(RegularCA):
    """Energy-conserving Life-like automaton (2-D Moore torus)."""

    energy = core.IntegerProperty(max_val=ENERGY_MASK)
    new_energy = core.IntegerProperty(max_val=ENERGY_MASK)
    birthing = core.IntegerProperty(max_val=1)
    rule = core.TotalisticRuleProperty(outer=True)
    rng = core.RandomProperty()
    death_speed = core.Parameter(default=1)
    full_treshold = core.Parameter(default=1)
    max_genes = core.Parameter(default=9)
    mutation_prob = core.Parameter(default=.0)

    def __init__(self, *args, legacy_coloring=True):
        """Pick the colouring (genome hue by default)."""
        self._legacy_coloring = legacy_coloring
        super().__init__(*args)

    def emit(self):
        """Decide whether to spread energy, and to which neighbour."""
        hood = len(self.buffers)
        limit = self.meta.full_treshold
        main = self.main
        main.new_energy = 1 * main.energy

        full_around = core.IntegerVariable()
        for i in range(hood):
            full_around += self.neighbors[i].main.energy >= limit

        is_full = main.energy >= limit
        is_empty = main.energy < limit
        not_sustained = xmath.int(
            xmath.int(main.rule.is_sustained(full_around)) == 0)
        born_here = main.rule.is_born(full_around)
        surplus = xmath.int(main.energy > limit)
        spread = (1 * is_full * (surplus | not_sustained)
                  + 1 * is_empty * (xmath.int(born_here) == 0))

        # scan the neighbours three times round from a random gate: first
        # pass birthing ones, second pass empty ones, the first fit wins
        portion = core.IntegerVariable()
        passed = core.IntegerVariable()
        passed *= 1
        start = core.IntegerVariable()
        start += xmath.int(main.rng.uniform * 8)
        target = core.IntegerVariable()
        for i in range(hood * 3):
            nbr = self.neighbors[i % 8].main
            in_window = (i >= start) * (i < start + hood * 2)
            birthing = nbr.birthing * (i < 8 + start)
            empty = (nbr.energy < limit) * (i >= 8 + start)
            portion *= 0
            portion += xmath.min(main.new_energy, self.meta.death_speed)
            portion *= spread * (empty + birthing) * (passed == 0)
            portion *= in_window
            target *= (passed != 0)
            target += (i % 8) * (passed == 0)
            passed += portion

        for i, buf in enumerate(self.buffers):
            chosen = i == target
            buf.energy = passed * chosen
            buf.rule = main.rule * chosen
        main.new_energy -= passed

    def absorb(self):
        """Collect the energy and the rules sent here."""
        hood = len(self.buffers)
        limit = self.meta.full_treshold
        received = core.IntegerVariable()
        full_around = core.IntegerVariable()
        for i in range(hood):
            received += self.neighbors[i].buffer.energy
            full_around += self.neighbors[i].main.new_energy >= limit
        self.main.energy = self.main.new_energy + received
        self.main.birthing = self.main.rule.is_born(full_around)
        self.main.birthing *= self.main.energy < limit

        offered = [core.IntegerVariable(name="genome%d" % i)
                   for i in range(hood)]
        for i in range(hood):
            sent = self.neighbors[i].buffer
            offered[i] += sent.rule * ((sent.energy > 0) * (sent.rule > 0))
        child = genome_crossover(
            self.main, self.main.rule.bit_width, *offered,
            max_genes=self.meta.max_genes,
            mutation_prob=self.meta.mutation_prob) * (self.main.energy < limit)
        self.main.rule = child + self.main.rule * (child == 0)

    @color_effects.MovingAverage
    def color(self):
        """Genome as hue, energy (up to the threshold) as brightness."""
        if self._legacy_coloring:
            red, green, blue = GenomeColor.modular(self.main.rule >> 1, 360)
        else:
            red, green, blue = GenomeColor.positional(
                self.main.rule, self.main.rule.bit_width)
        shade = xmath.min(self.main.energy, self.meta.full_treshold)
        shade = shade * 255 / self.meta.full_treshold
        return (xmath.int(red * shade), xmath.int(green * shade),
                xmath.int(blue * shade))


class ConservedLife4D(ConservedLife):
    """The same rules on a 4-D von Neumann torus (8 neighbours)."""

    class Topology(ConservedLife.Topology):
        """4-D von Neumann neighbourhood."""

        dimensions = 4
        neighborhood = core.VonNeumannNeighborhood()


def _rules(word, count, mask):
    """Reproducible random rules (the reference draws them unseeded)."""
    draw = random.Random(word)
    return [draw.randint(0, 2 ** 18 - 1) & mask for _ in range(count + 1)]


class BigBangExperiment(RegularExperiment):
    """640x360: 18 squares of full cells with random rules."""

    death_speed = 1
    full_treshold = 16
    max_genes = 13
    fade_in = 255
    fade_out = 255
    smooth_ratio = 1
    rule_mask = LifeLike.golly2int("B12345678/S12345678")
    _drawn = _rules("big bang", 18, rule_mask)
    seed = seeds.patterns.PrimordialSoup(vals={
        "energy": 0, "new_energy": 0, "rule": _drawn[0],
        "rng": RandInt(0, 2 ** 16 - 1)})
    for _i in range(18):
        seed += seeds.patterns.BigBang(
            pos=(30 + (_i % 6) * 100, 40 + (_i // 6) * 100), size=(80, 80),
            vals={"energy": 16 * RandInt(0, 1),
                  "new_energy": 16 * RandInt(0, 1),
                  "rule": _drawn[_i + 1], "rng": RandInt(0, 2 ** 16 - 1)})


class BigBangExperiment4D(RegularExperiment):
    """640x360x3x3 with 273-energy squares (the reference's 4-D run)."""

    size = (640, 360, 3, 3)
    death_speed = 1
    full_treshold = 16
    max_genes = 5
    fade_in = 1
    fade_out = 1
    smooth_ratio = 10
    rule_back = LifeLike.golly2int("B12/S124")
    rule_area = LifeLike.golly2int("B12/S124")
    seed = seeds.patterns.PrimordialSoup(vals={
        "energy": 2, "new_energy": 2, "rule": rule_back,
        "rng": RandInt(0, 2 ** 16 - 1)})
    for _i in range(18):
        seed += seeds.patterns.BigBang(
            pos=(30 + (_i % 6) * 100, 40 + (_i // 6) * 100, 0, 0),
            size=(80, 80, 1, 1),
            vals={"energy": 273 * RandInt(0, 1),
                  "new_energy": 273 * RandInt(0, 1),
                  "rule": rule_area, "rng": RandInt(0, 2 ** 16 - 1)})
