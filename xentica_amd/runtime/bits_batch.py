"""
Batches of Life-like 2-D tori on the bit-sliced batch kernel.

:class:`BitsBatch` runs hundreds of independent automata of uneven sizes,
each with its own rule, in one launch per ``depth`` steps
(``xen_life_batch_launch``, ``csrc/life_batch.hip``), a multi-step run
captured in a HIP graph.  Every instance is a model whose generated code
:class:`runtime.hostc.LifeLikeProbe` proves to be ``next = born(count) |
sustain(count) & state`` on a 1-bit Moore torus (the reference's
``LifelikeCA``, ``examples/game_of_life.py:183-203``), with its own
parameter values; any width of at least 32 cells and any height.

:func:`plan_lanes` lays the work out.  Each instance is split into column
tiles of at most 62 stored words and row strips of about ``strip_rows``
rows; a tile-strip takes its words plus two halo lanes, and tile-strips
are packed into the 64 lanes of a wave, the tallest first, so that narrow
instances share waves instead of leaving most lanes idle.
"""
import ctypes

import numpy as np

from xentica_amd.core.exceptions import XenticaException
from xentica_amd.runtime import hip
from xentica_amd.runtime.hostc import LifeLikeProbe

__all__ = ['plan_lanes', 'balanced_strips', 'BitsBatch', 'autotune',
           'MAX_TILE_WORDS', 'WAVE']

#: Lanes of a wavefront, and stored words of one column tile (two lanes of
#: a wave's 64 are the tile's halo).
WAVE = 64
MAX_TILE_WORDS = WAVE - 2

_vp = ctypes.c_void_p
_SIGNATURES = {
    'xen_life_batch_launch': [_vp, _vp, ctypes.c_longlong, _vp, _vp, _vp,
                              ctypes.c_int, ctypes.c_int, _vp],
    'xen_bits_pack_rows': [_vp, _vp, ctypes.c_int, ctypes.c_int,
                           ctypes.c_int, _vp],
    'xen_bits_expand_rows': [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                             ctypes.c_int, _vp, _vp, _vp],
}
_BOUND = False


def _lib():
    global _BOUND  # pylint: disable=global-statement
    lib = hip.library()
    if not _BOUND:
        for name, argtypes in _SIGNATURES.items():
            func = getattr(lib, name)
            func.argtypes = argtypes
            func.restype = ctypes.c_int
        _BOUND = True
    return lib


def _stream(stream):
    return stream.handle if stream is not None else None


def pitch_of(width):
    """Stored 32-bit words per row."""
    return (int(width) + 31) // 32


def layout(sizes):
    """Word offset of each instance and the batch's total words."""
    bases, total = [], 0
    for width, height in sizes:
        bases.append(total)
        total += pitch_of(width) * int(height)
    return bases, total


def _pack(sizes, strip_rows):
    """The waves of a plan: ``[free lanes, rows, [items]]`` each, an item
    being ``(rows, lanes, instance, first word, end word, first row)``."""
    items = []
    for inst, (width, height) in enumerate(sizes):
        width, height = int(width), int(height)
        if width < 32 or height < 1:
            raise XenticaException(
                "bit-sliced batch: instance %d is %dx%d (width must be at "
                "least 32)" % (inst, width, height))
        pitch = pitch_of(width)
        tiles = -(-pitch // MAX_TILE_WORDS)
        strips = max(1, -(-height // strip_rows))
        for tile in range(tiles):
            col_lo = pitch * tile // tiles
            col_hi = pitch * (tile + 1) // tiles
            for strip in range(strips):
                row_lo = height * strip // strips
                rows = height * (strip + 1) // strips - row_lo
                items.append((rows, col_hi - col_lo + 2, inst, col_lo,
                              col_hi, row_lo))
    # tallest strips first, then the widest: each goes into the first wave
    # with room for it, whose rows are at least its own
    items.sort(key=lambda item: (-item[0], -item[1], item[2], item[3],
                                 item[5]))
    waves = []          # [free lanes, rows, [items]]
    open_waves = []
    for item in items:
        rows, need = item[0], item[1]
        for wave in open_waves:
            if wave[0] >= need:
                break
        else:
            wave = [WAVE, rows, []]
            waves.append(wave)
            open_waves.append(wave)
        wave[0] -= need
        wave[2].append(item)
        if wave[0] < 3:
            open_waves.remove(wave)
    return waves


def balanced_strips(sizes, simds, per_simd=(2, 3), lowest=8, highest=512):
    """
    Strip heights whose plans come closest to, without exceeding, a whole
    number of waves per SIMD: for each count in ``per_simd``, the lowest
    height in ``[lowest, highest]`` whose plan has at most ``count * simds``
    waves (none when even the lowest height stays under that count).  A VALU-bound launch whose waves are one short of filling the
    SIMDs evenly runs as long as one that fills them, so these heights are
    where the halo rows are fewest for a given finishing time.
    """
    out = []
    for count in per_simd:
        limit = count * simds
        low, high = lowest, highest
        if len(_pack(sizes, high)) > limit or \
                len(_pack(sizes, low)) <= limit:
            continue  # no height in range reaches (or stays under) it
        while low < high:
            mid = (low + high) // 2
            if len(_pack(sizes, mid)) <= limit:
                high = mid
            else:
                low = mid + 1
        out.append(low)
    return sorted(set(out))


def plan_lanes(sizes, strip_rows=48):
    """
    The lane descriptors of one launch over instances of ``sizes``.

    Returns ``(lanes, wave_rows)``: ``lanes`` is int32 ``[waves, 64, 4]``
    (instance or -1, extended word column, first output row, output rows
    or 0 for lanes that store nothing) and ``wave_rows`` the most output
    rows of each wave.  Every stored word of every instance is the output
    of exactly one lane.
    """
    waves = _pack(sizes, strip_rows)
    lanes = np.zeros((len(waves), WAVE, 4), dtype=np.int32)
    lanes[:, :, 0] = -1
    wave_rows = np.zeros(len(waves), dtype=np.int32)
    for index, (_, rows, members) in enumerate(waves):
        wave_rows[index] = rows
        lane = 0
        for item_rows, need, inst, col_lo, col_hi, row_lo in members:
            for k in range(need):
                col = col_lo - 1 + k
                stores = col_lo <= col < col_hi
                lanes[index, lane] = (inst, col, row_lo,
                                      item_rows if stores else 0)
                lane += 1
    return lanes, wave_rows


class BitsBatch:
    """
    Independent Life-like 2-D tori of one or more model classes advanced
    together on the bit-sliced batch kernel.

    :param models: models whose step is a Life-like rule on a 1-bit Moore
        torus for their own parameter values (``model.<param> = value``)
    :param depth: steps per launch (1..8)
    :param strip_rows: about how many rows one lane's strip has
    """

    def __init__(self, models, depth=8, strip_rows=48, stream=None):
        self.models = list(models)
        if not self.models:
            raise XenticaException("empty batch")
        self.depth, self.strip_rows = int(depth), int(strip_rows)
        if not 1 <= self.depth <= 8:
            raise XenticaException("bit-sliced batch: depth must be 1..8")
        self.stream = stream or hip.Stream()
        self._probes = {}
        for model in self.models:
            if len(model.size) != 2:
                raise XenticaException("bit-sliced batch: 2-D models only")
            self._probe(model)
        self.sizes = [tuple(int(v) for v in model.size)
                      for model in self.models]
        self.bases, total_words = layout(self.sizes)
        self.total_bytes = 4 * total_words
        if self.total_bytes >= 1 << 31:
            raise XenticaException("bit-sliced batch: more than 2 GiB")
        lanes, wave_rows = plan_lanes(self.sizes, self.strip_rows)
        self.n_waves = len(wave_rows)
        self.lanes = lanes
        self.wave_rows = wave_rows
        self._lanes_buf = hip.Buffer(lanes.nbytes)
        self._lanes_buf.upload(lanes)
        self._rows_buf = hip.Buffer(wave_rows.nbytes)
        self._rows_buf.upload(wave_rows)
        self.table = np.zeros((len(self.models), 8), dtype=np.int32)
        for idx, ((width, height), base) in enumerate(zip(self.sizes,
                                                          self.bases)):
            self.table[idx, :4] = (base, pitch_of(width), width, height)
        self._table_buf = hip.Buffer(self.table.nbytes)
        self._uploaded = None
        self.state = [hip.Buffer(self.total_bytes),
                      hip.Buffer(self.total_bytes)]
        self.cur = 0
        self._graphs = {}
        self.gather()

    # -- rules --------------------------------------------------------------

    def _probe(self, model):
        cls = type(model)
        if cls not in self._probes:
            self._probes[cls] = LifeLikeProbe(model)
        return self._probes[cls]

    def rules(self):
        """``(born, sustain, colours)`` of every instance for its current
        parameter values."""
        out = []
        for idx, model in enumerate(self.models):
            probe = self._probe(model)
            own = getattr(model, "_param_values", {})
            values = tuple(own.get(p.name, p.value) for p in probe.params)
            rule = probe.rule(values)
            if rule is None:
                raise XenticaException(
                    "bit-sliced batch: instance %d is not a Life-like rule "
                    "for its parameter values %r" % (idx, values))
            out.append(rule)
        return out

    def _refresh(self):
        """Upload the rule table when parameter values changed (ordered on
        the batch's stream, after the launches that read the old one)."""
        rules = self.rules()
        key = tuple((born, sustain) for born, sustain, _ in rules)
        if key != self._uploaded:
            for idx, (born, sustain) in enumerate(key):
                self.table[idx, 4:6] = (born, sustain)
            self._table_buf.upload(self.table, self.stream)
            self._uploaded = key
        return rules

    # -- state in and out ------------------------------------------------------

    def gather(self):
        """Pack every model's main plane into the batch (the models' queued
        steps first); their stored colours must be ones the colour proof
        covers."""
        for model in self.models:
            model._engine.flush()
        hip.device().synchronize()
        lib = _lib()
        for idx, (model, base) in enumerate(zip(self.models, self.bases)):
            probe = self._probe(model)
            bounds = model.gpu.arrays.colors.bounds
            if probe.plan and (bounds is None or any(
                    not (low <= lo and hi <= high)
                    for (lo, hi), (low, high) in zip(bounds, probe.plan))):
                raise XenticaException(
                    "bit-sliced batch: instance %d holds colours the "
                    "colour proof does not cover" % idx)
            width, height = self.sizes[idx]
            hip.check(lib.xen_bits_pack_rows(
                model.gpu.arrays.cells.plane_ptr(model.cells_num),
                self.state[self.cur].ptr + 4 * base, width, height,
                pitch_of(width), _stream(self.stream)), "xen_bits_pack_rows")
        self.stream.synchronize()

    def scatter(self):
        """Expand the batch into every model's main plane and colours."""
        lib = _lib()
        rules = self.rules()
        for idx, (model, base) in enumerate(zip(self.models, self.bases)):
            width, height = self.sizes[idx]
            dead, alive = rules[idx][2]
            arrays = model.gpu.arrays
            hip.check(lib.xen_bits_expand_rows(
                self.state[self.cur].ptr + 4 * base,
                arrays.cells.plane_ptr(model.cells_num),
                arrays.colors.device_ptr(),
                width, height, pitch_of(width),
                (ctypes.c_int32 * 3)(*dead), (ctypes.c_int32 * 3)(*alive),
                _stream(self.stream)), "xen_bits_expand_rows")
            arrays.colors.bounds = tuple(
                (min(a, b), max(a, b)) for a, b in zip(dead, alive))
            engine = model._engine
            if hasattr(engine, "invalidate"):
                engine.invalidate()
        self.stream.synchronize()

    def save(self, filename):
        """Write every instance's state in the reference's checkpoint
        format (``xentica/core/base.py:755-763``): a pickled
        ``{"instances": [{"cells", "colors", "random"}, ...]}`` in the
        order of :attr:`models`; each entry is what that model's own
        ``save()`` writes."""
        from xentica_amd.runtime import checkpoint
        self.synchronize()
        self.scatter()
        checkpoint.write(filename, {"instances": [
            {"cells": model.gpu.arrays.cells.get(),
             "colors": model.gpu.arrays.colors.get(),
             "random": model.random} for model in self.models]})

    def load(self, filename):
        """Read a checkpoint written by :meth:`save` into the models and
        the batch (``xentica/core/base.py:765-771`` per instance)."""
        from xentica_amd.runtime import checkpoint
        states = checkpoint.read(filename)["instances"]
        if len(states) != len(self.models):
            raise XenticaException("checkpoint holds %d instances, the batch "
                                   "%d" % (len(states), len(self.models)))
        self.synchronize()
        for model, state in zip(self.models, states):
            model._engine.load_state(state["cells"], state["colors"])
            model.random.load(state["random"])
        self.gather()

    # -- stepping ------------------------------------------------------------------

    def _launch(self, steps):
        hip.check(_lib().xen_life_batch_launch(
            self.state[self.cur].ptr, self.state[1 - self.cur].ptr,
            self.total_bytes, self._table_buf.ptr, self._lanes_buf.ptr,
            self._rows_buf.ptr, self.n_waves, steps, _stream(self.stream)),
            "xen_life_batch_launch")
        self.cur = 1 - self.cur

    def launch_plan(self, steps):
        """Steps per launch of a run of ``steps`` steps."""
        depth = self.depth
        return [depth] * (steps // depth) + (
            [steps % depth] if steps % depth else [])

    def run(self, steps, graph=True):
        """Advance every instance ``steps`` steps (asynchronous)."""
        self._refresh()
        plan = self.launch_plan(steps)
        if not plan:
            return
        if not graph:
            for count in plan:
                self._launch(count)
            return
        key = (tuple(plan), self.cur)
        if key not in self._graphs:
            start = self.cur
            hip.Graph.begin(self.stream)
            try:
                for count in plan:
                    self._launch(count)
            finally:
                graph_obj = hip.Graph.end(self.stream)
            self._graphs[key] = (graph_obj, self.cur)
            self.cur = start
        graph_obj, end = self._graphs[key]
        graph_obj.launch(self.stream)
        self.cur = end

    def synchronize(self):
        """Wait for the batch's stream."""
        self.stream.synchronize()

    def lane_use(self):
        """Fraction of the launched lanes that store a word."""
        return float((self.lanes[:, :, 3] > 0).sum()) / self.lanes[:, :, 0].size


#: Temporal depths and fixed strip heights the autotuner tries (beside
#: the heights :func:`balanced_strips` finds for the batch).
TUNE_DEPTHS = (4, 6, 8)
TUNE_STRIPS = (32, 48)


def autotune(models, trial_steps=48, depths=TUNE_DEPTHS,
             strip_rows=None, stream=None, log=None):
    """
    Pick the temporal depth and the strip height of a bit-sliced batch by
    measurement: every pair runs once to warm up, then ``trial_steps``
    steps (whole launches, graph-captured) between HIP events, from the
    models' states (which are not advanced).  The heights tried are
    ``strip_rows``, or by default :data:`TUNE_STRIPS` and the heights that
    bring the plan just under two and three waves per SIMD of this device.
    Returns ``(batch, timings)`` with the fastest :class:`BitsBatch` ready
    to run from the models' states and ``{(depth, strip_rows): cell-updates
    per second}``.
    """
    models = list(models)
    if strip_rows is None:
        simds = 4 * hip.device().compute_units
        strip_rows = sorted(set(TUNE_STRIPS) | set(balanced_strips(
            [model.size for model in models], simds)))
    stream = stream or hip.Stream()
    cells = sum(model.cells_num for model in models)
    timings, best = {}, None
    for depth in depths:
        for rows in strip_rows:
            batch = BitsBatch(models, depth=depth, strip_rows=rows,
                              stream=stream)
            steps = -(-trial_steps // depth) * depth
            batch.run(steps)
            start, stop = hip.Event(), hip.Event()
            start.record(stream)
            batch.run(steps)
            stop.record(stream)
            stop.synchronize()
            rate = cells * steps / (start.elapsed_ms(stop) / 1e3)
            timings[(depth, rows)] = rate
            if log:
                log("bits batch depth %d strip rows %d (%d waves): %.4g "
                    "cell-updates/s" % (depth, rows, batch.n_waves, rate))
            if best is None or rate > best[0]:
                best = (rate, depth, rows)
            del batch
    _, depth, rows = best
    return BitsBatch(models, depth=depth, strip_rows=rows,
                     stream=stream), timings
