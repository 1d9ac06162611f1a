"""
Benchmark and parity workloads, sized as the project's measurement plan.

* :class:`Life512Experiment` -- Life B3/S23 on a 512x512 torus, 50 %
  random state from seed 42, used for per-step hash parity.
* :class:`Life16kExperiment` -- the same at 16384x16384 (the headline
  throughput workload, run by the bit-packed engine).
* :class:`LifeLikeRule` and :func:`batch_experiments` -- the batch of 256
  Life-like automata of uneven sizes, each with its own rule and density.
"""
from xentica import core
from xentica import seeds
from xentica.core import color_effects


class Life512Experiment(core.Experiment):
    """512x512, ``RandInt(0, 1)`` everywhere, word 42."""

    word = 42
    size = (512, 512)
    zoom = 1
    pos = [0, 0]
    seed = seeds.patterns.PrimordialSoup(
        vals={"state": seeds.random.RandInt(0, 1)})


class Life16kExperiment(Life512Experiment):
    """16384x16384, ``RandInt(0, 1)`` everywhere, word 42."""

    size = (16384, 16384)


class LifeLikeRule(core.CellularAutomaton):
    """
    Life-like automaton whose rule is an interactive parameter.

    Birth counts in bits 0-8 and survival counts in bits 9-17 of ``rule``
    (the encoding of ``LifeLike.golly2int``); ``next = born | sustained &
    state`` as in the reference's ``LifelikeCA`` (without its rule switch).
    """

    state = core.IntegerProperty(max_val=1)
    rule = core.Parameter(default=6152, interactive=True)

    class Topology:
        """2-D Moore torus."""

        dimensions = 2
        lattice = core.OrthogonalLattice()
        neighborhood = core.MooreNeighborhood()
        border = core.TorusBorder()

    def emit(self):
        """Every buffer carries the state."""
        for buf in self.buffers:
            buf.state = self.main.state

    def absorb(self):
        """Apply the rule to the live-neighbour count."""
        count = core.IntegerVariable()
        for i in range(len(self.buffers)):
            count += self.neighbors[i].buffer.state
        is_born = (self.rule >> count) & 1
        is_sustain = (self.rule >> 9 >> count) & 1
        self.main.state = is_born | is_sustain & self.main.state

    @color_effects.MovingAverage
    def color(self):
        """Live cells white."""
        val = self.main.state * 255
        return (val, val, val)


def batch_experiments(count=256, seed=42, side_min=64, side_max=2048):
    """
    Experiments of the batched workload: ``count`` square tori.

    Sides are ``round(exp(U(ln side_min, ln side_max)))``, birth and
    survival masks uniform 9-bit, density uniform in [0.1, 0.6] quantised
    to d/1024 (seeded as ``(RandInt(0, 1023) + d) >> 10``), word
    ``seed + index``; drawn from ``numpy.random.default_rng(seed)``.
    """
    import numpy as np
    rng = np.random.default_rng(seed)
    sides = np.round(np.exp(rng.uniform(np.log(side_min), np.log(side_max),
                                        count))).astype(int)
    born = rng.integers(0, 512, count)
    survive = rng.integers(0, 512, count)
    density = rng.uniform(0.1, 0.6, count)
    experiments = []
    for idx in range(count):
        level = int(round(density[idx] * 1024))
        attrs = {
            "word": seed + idx,
            "size": (int(sides[idx]), int(sides[idx])),
            "zoom": 1,
            "pos": [0, 0],
            "rule": int(born[idx]) | (int(survive[idx]) << 9),
            "seed": seeds.patterns.PrimordialSoup(vals={
                "state": (seeds.random.RandInt(0, 1023) + level) >> 10}),
        }
        experiments.append(type("BatchExperiment%d" % idx,
                                (core.Experiment,), attrs))
    return experiments


class LifeLike3D(core.CellularAutomaton):
    """
    Life-like rule on a 3-D Moore torus (26 offsets, the reference's list).

    ``born``/``sustain`` are count masks (bit c: c counted cells); the
    count follows the reference's 3-D ``MooreNeighborhood``, which keeps the
    centre and drops (1, 1, 1).
    """

    state = core.IntegerProperty(max_val=1)
    born = core.Parameter(default=1 << 5, interactive=True)
    sustain = core.Parameter(default=(1 << 4) | (1 << 5), interactive=True)

    class Topology:
        """3-D Moore torus."""

        dimensions = 3
        lattice = core.OrthogonalLattice()
        neighborhood = core.MooreNeighborhood()
        border = core.TorusBorder()

    def emit(self):
        """Every buffer carries the state."""
        for buf in self.buffers:
            buf.state = self.main.state

    def absorb(self):
        """Apply the masks to the count."""
        count = core.IntegerVariable()
        for i in range(len(self.buffers)):
            count += self.neighbors[i].buffer.state
        is_born = (self.born >> count) & 1
        is_sustain = (self.sustain >> count) & 1
        self.main.state = is_born | is_sustain & self.main.state

    @color_effects.MovingAverage
    def color(self):
        """Live cells white (projected by the renderer)."""
        val = self.main.state * 255
        return (val, val, val)


class Life3DExperiment(core.Experiment):
    """B5/S45 at 10 % density (``(RandInt(0, 9) + 7) >> 4``), word 42."""

    word = 42
    size = (1024, 1024, 1024)
    zoom = 1
    pos = [0, 0, 0]
    born = 1 << 5
    sustain = (1 << 4) | (1 << 5)
    seed = seeds.patterns.PrimordialSoup(
        vals={"state": (seeds.random.RandInt(0, 9) + 7) >> 4})


def life3d_cells(size=(1024, 1024, 1024), word=42, chunk=1 << 26):
    """
    Step-0 cells of :class:`Life3DExperiment` at ``size``, drawn in chunks.

    Identical to ``PrimordialSoup({'state': (RandInt(0, 9) + 7) >> 4})``:
    the legacy ``RandomState.randint`` stream continues across calls, so
    chunked draws equal one draw of the whole field.
    """
    import numpy as np
    from xentica_amd.seeds.random import LocalRandom
    stream = LocalRandom(word).numpy
    n = int(np.prod(size))
    cells = np.empty(n, dtype=np.uint8)
    for start in range(0, n, chunk):
        count = min(chunk, n - start)
        cells[start:start + count] = (stream.randint(0, 10, count) + 7) >> 4
    return cells
