"""
The ``bits`` engine: Life-like models stepped on the bit-sliced kernels.

A model qualifies when :class:`runtime.hostc.LifeLikeProbe` shows, from
its own generated code and for the parameter values in force, that its
step is ``next = born(count) | sustain(count) & state`` on a 1-bit state
with a colour that is a function of the new state (the reference's
``GameOfLife`` and ``LifelikeCA``, ``examples/game_of_life.py:15-90,
183-203``, and any model written the same way).  The state is then held
packed, 32 cells per 32-bit word, and advanced by ``xen_life_bits_run``
(2-D, ``depth`` steps per launch in registers) or ``xen_life3d_bits_step``
(3-D, the reference's 26-offset Moore list); ``step()`` defers steps
until a launch is full, a parameter changes or the state is read.

The reference's arrays stay the model's interface: reading
``gpu.arrays.cells`` or ``gpu.arrays.colors`` (``.get()``/``.set()``),
``render()``, ``save()`` and batches first expand the packed state into
the main plane and the colours in one pass (``xen_bits_expand``), and the
next step packs the main plane again.  Steps whose parameter values make
the model something other than a Life-like rule, or that start from
stored colours the colour proof does not cover, run on the fused engine
(or the plain one) instead, from the same arrays.
"""
import numpy as np

from xentica_amd.core.exceptions import XenticaException
from xentica_amd.runtime import hip, kernels
from xentica_amd.runtime.hostc import LifeLikeProbe

__all__ = ['BitsEngine', 'eligible_size']


def eligible_size(size):
    """Whether the bit-sliced kernels take a field of ``size``: 2-D fields
    at least 32 cells wide (widths that are not a multiple of 32 run on
    the batch kernel's padded rows), 3-D ones whose width / 32 divides 64."""
    if len(size) == 2:
        width, height = size
        return width >= 32 and height > 0 and \
            4 * ((width + 31) // 32) * height < 1 << 31
    if len(size) == 3:
        width, height, depth = size
        return (width % 32 == 0 and width > 0 and 64 % (width // 32) == 0
                and height > 0 and depth > 0
                and width * height * depth // 8 < 1 << 31)
    return False


class BitsEngine:
    """Single-model engine on the bit-sliced Life-like kernels."""

    name = "bits"

    def __init__(self, model, depth=8, strips=0):
        self.model = model
        if not eligible_size(model.size):
            raise XenticaException(
                "bits engine: the field %s does not fit the bit-sliced "
                "kernels (2-D: at least 32 cells wide; 3-D: width a "
                "multiple of 32 with width / 32 dividing 64)"
                % (model.size, ))
        self.probe = LifeLikeProbe(model)
        if self.probe.rule(self.probe.values()) is None:
            raise XenticaException(
                "bits engine: the model is not a Life-like rule for its "
                "parameter values")
        #: steps per launch (2-D) and row strips (0: the library's choice)
        self.depth, self.strips = int(depth), int(strips)
        self.dims = len(model.size)
        self.n = model.cells_num
        #: 2-D widths that are not a multiple of 32: rows padded to whole
        #: words, stepped by the batch kernel as a batch of one
        self.rows = _PaddedRows(model.size, depth) if (
            self.dims == 2 and model.size[0] % 32) else None
        self._bufs = None
        self._cur = 0
        self._packed = False        # the packed buffer holds the state
        self._stale = False         # cells and colours lag the packed state
        self._pending = 0
        self._pending_rule = None
        self._fallback = None
        self._in_fallback = False
        self._hook()

    def _hook(self):
        arrays = self.model.gpu.arrays
        arrays.cells.before_read = self.flush
        arrays.colors.before_read = self.flush

    # -- the engine interface -------------------------------------------

    def launch_elementwise(self, kernel, args, count):
        """Grid-stride launch (render, through the plain kernels)."""
        from xentica_amd.runtime.engines import elementwise_grid, BLOCK
        kernel(*args, grid=(elementwise_grid(count), 1, 1),
               block=(BLOCK, 1, 1))

    def step(self):
        """Queue one step with the current parameter values."""
        rule = self.probe.rule(self.probe.values())
        if rule is None or (not self._packed and not self._colors_admitted()):
            self._fallback_step()
            return
        if self._in_fallback:
            self._fallback.flush()
            self._in_fallback = False
        # the probe returns the same rule object for the same values
        if self._pending and rule is not self._pending_rule and \
                rule != self._pending_rule:
            self._launch()
        self._pending_rule = rule
        self._pending += 1
        if self.dims == 3 or self._pending >= self.depth:
            self._launch()

    def finish(self):
        """Launch the queued steps without expanding the state (the
        arrays are brought up to date by the next read)."""
        if self._pending:
            self._launch()
        if self._in_fallback:
            self._fallback.flush()

    def flush(self):
        """Run the queued steps; leave the state in the reference's arrays."""
        self.finish()
        if self._stale:
            self._expand()

    def load_state(self, cells, colors):
        """Replace the state (queued steps are applied first)."""
        self.flush()
        arrays = self.model.gpu.arrays
        arrays.cells.set(cells)
        arrays.colors.set(colors)
        self._packed = False

    def render_frame(self, render_params, pixels):
        """
        Project the packed state into the model's image with the model's
        own render body (``renderers.py:220-270``), its colour load replaced
        by the colour of the cell's bit: the colours are never expanded.
        Returns False (the caller renders from the arrays) when the packed
        state is not the current one or the render body is not the plain
        renderer's.
        """
        if self._pending:
            self._launch()
        if not (self._packed and self._stale) or self.rows is not None:
            return False
        function = self._render_function()
        if function is None:
            return False
        model = self.model
        _, _, (dead, alive) = self._pending_rule
        args = model.renderer.get_args_vals(model)
        args[0] = np.uint64(self._bufs[self._cur].ptr)
        args += list(render_params)
        args += [np.int32(v) for v in dead + alive]
        args.append(np.int64(pixels))
        self.launch_elementwise(function, args, pixels)
        return True

    _COLOUR_LOAD = "int3 c = col[ii];\n"

    def _render_function(self):
        if getattr(self, "_render_fn", False) is not False:
            return self._render_fn
        self._render_fn = None
        model = self.model
        body = model.kernel_bodies.get('render', "")
        if body.count(self._COLOUR_LOAD) != 1 or model.renderer.args[0] != \
                ("int3", "*col"):
            return None
        from xentica_amd.core.base import _elementwise_kernel
        body = body.replace(self._COLOUR_LOAD, (
            "int3 c = ((_bits[(unsigned) ii >> 5] >> (ii & 31)) & 1u) ?\n"
            "    make_int3(_a0, _a1, _a2) : make_int3(_d0, _d1, _d2);\n"))
        args = [("const unsigned", "*_bits")] + list(model.renderer.args[1:])
        args += [(p.ctype, p.name) for p in model._render_params]
        args += [("int", name) for name in
                 ("_d0", "_d1", "_d2", "_a0", "_a1", "_a2")]
        source = model.build_defines() + _elementwise_kernel(
            "render_bits", args, body)
        module = hip.Module(hip.compile_source(
            source, "bsca_render_bits.hip",
            hip.compile_options(model.fp_contract)))
        self._render_module = module
        self._render_fn = module.get_function("render_bits")
        return self._render_fn

    def invalidate(self):
        """The arrays were written directly (a batch's ``scatter()``): the
        next step packs them again."""
        self._pending = 0
        self._packed = False
        self._stale = False

    # -- internals --------------------------------------------------------

    def _colors_admitted(self):
        plan = self.probe.plan
        if plan == ():
            return True
        bounds = self.model.gpu.arrays.colors.bounds
        return bounds is not None and all(
            low <= lo and hi <= high
            for (lo, hi), (low, high) in zip(bounds, plan))

    def _fallback_step(self):
        """One step on the fused (or plain) engine from the arrays."""
        self.flush()
        self._packed = False
        if self._fallback is None:
            from xentica_amd.runtime import engines
            self._fallback = engines.create(self.model, "fused_or_plain")
            self._hook()
        self._in_fallback = True
        self._fallback.step()

    def _buffers(self):
        if self._bufs is None:
            nbytes = self.rows.nbytes if self.rows else self.n // 8
            self._bufs = [hip.Buffer(nbytes), hip.Buffer(nbytes)]
        return self._bufs

    def _pack(self):
        bufs = self._buffers()
        plane = self.model.gpu.arrays.cells.plane_ptr(self.n)
        if self.rows:
            self.rows.pack(plane, bufs[self._cur].ptr)
        else:
            kernels.bits_pack(plane, bufs[self._cur].ptr, self.n)
        self._packed = True

    def _launch(self):
        steps, (born, sustain, _) = self._pending, self._pending_rule
        self._pending = 0
        if not self._packed:
            self._pack()
        bufs = self._buffers()
        if self.rows:
            for count in self.rows.launch_plan(steps):
                self.rows.launch(bufs[self._cur].ptr,
                                 bufs[1 - self._cur].ptr, count, born,
                                 sustain)
                self._cur = 1 - self._cur
        elif self.dims == 2:
            width, height = self.model.size
            in_b = kernels.life_bits_run(
                bufs[self._cur].ptr, bufs[1 - self._cur].ptr, width // 32,
                height, born, sustain, steps, self.depth, self.strips)
            if in_b:
                self._cur = 1 - self._cur
        else:
            for _ in range(steps):
                kernels.life3d_bits_step(
                    bufs[self._cur].ptr, bufs[1 - self._cur].ptr,
                    self.model.size, born, sustain)
                self._cur = 1 - self._cur
        self._stale = True

    def _expand(self):
        """The packed state into the main plane and the colours."""
        arrays = self.model.gpu.arrays
        _, _, (dead, alive) = self._pending_rule
        if self.rows:
            self.rows.expand(self._bufs[self._cur].ptr,
                             arrays.cells.plane_ptr(self.n),
                             arrays.colors.device_ptr(), dead, alive)
        else:
            kernels.bits_expand(self._bufs[self._cur].ptr,
                                arrays.cells.plane_ptr(self.n),
                                arrays.colors.device_ptr(), self.n, dead,
                                alive)
        arrays.colors.bounds = tuple(
            (min(a, b), max(a, b)) for a, b in zip(dead, alive))
        self._stale = False
        # the arrays may be written next (``set()``): pack them again
        # before the next step
        self._packed = False

    @property
    def state_ptr(self):
        """Device address of the packed state (launch queued steps first;
        rows padded to whole words when the width is not a multiple of
        32)."""
        self.finish()
        if not self._packed:
            self._pack()
        return self._bufs[self._cur].ptr

    def crc32(self):
        """``binascii.crc32`` of the main plane, from the packed state (from
        the expanded plane for padded rows)."""
        if self.rows:
            self.flush()
            return kernels.crc32_u8(
                self.model.gpu.arrays.cells.plane_ptr(self.n), self.n)
        return kernels.crc32_bits(self.state_ptr, self.n)


class _PaddedRows:
    """
    A 2-D torus whose width is not a multiple of 32, packed as rows of
    ``ceil(width / 32)`` words (the bits past the width zero) and stepped
    by the batch kernel (``xen_life_batch_launch``) as a batch of one
    instance: its lanes read 32-cell windows that run over the row's end
    into the row's start, so the torus wrap is exact for any width.  All
    calls go to the null stream, in order with the engine's other work.
    """

    def __init__(self, size, depth):
        from xentica_amd.runtime import bits_batch
        self.width, self.height = (int(v) for v in size)
        self.pitch = bits_batch.pitch_of(self.width)
        self.nbytes = 4 * self.pitch * self.height
        self.depth = int(depth)
        self._lib = None
        self._rule = None

    def _setup(self):
        """The lane plan and the instance table on the device (on first
        use, so that models build without a GPU)."""
        if self._lib is not None:
            return
        from xentica_amd.runtime import bits_batch
        size = (self.width, self.height)
        simds = 4 * hip.device().compute_units
        lanes, wave_rows = bits_batch.plan_lanes(
            [size], self.strip_rows(simds))
        self.n_waves = len(wave_rows)
        self._lanes = hip.Buffer(lanes.nbytes)
        self._lanes.upload(lanes)
        self._rows = hip.Buffer(wave_rows.nbytes)
        self._rows.upload(wave_rows)
        self._table = np.zeros((1, 8), dtype=np.int32)
        self._table[0, :4] = (0, self.pitch, self.width, self.height)
        self._table_buf = hip.Buffer(self._table.nbytes)
        self._lib = bits_batch._lib()

    def strip_rows(self, simds):
        """Rows per strip by the cost model of the 2-D kernel's default
        (``auto_strips`` in ``csrc/life_bits.hip``): for one, two and three
        waves per SIMD, (waves per SIMD) x (strip rows + 2K) over the share
        of VALU slots that many waves keep busy (0.55, 0.85, 1)."""
        from xentica_amd.runtime import bits_batch
        tiles = -(-self.pitch // bits_batch.MAX_TILE_WORDS)
        best = None
        for waves, issue in ((1, 0.55), (2, 0.85), (3, 1.0)):
            strips = max(1, min(self.height, (waves * simds) // tiles))
            per_simd = max(1.0, strips * tiles / simds)
            time = per_simd * (self.height / strips + 2 * self.depth) / issue
            if best is None or time < best[0]:
                best = (time, strips)
        return -(-self.height // best[1])

    def launch_plan(self, steps):
        """Steps per launch of ``steps`` queued steps."""
        return [self.depth] * (steps // self.depth) + (
            [steps % self.depth] if steps % self.depth else [])

    def pack(self, plane, bits):
        """uint8 main plane -> padded packed rows."""
        self._setup()
        hip.check(self._lib.xen_bits_pack_rows(
            plane, bits, self.width, self.height, self.pitch, None),
            "xen_bits_pack_rows")

    def expand(self, bits, plane, colors, dead, alive):
        """Padded packed rows -> uint8 main plane and colours by state."""
        import ctypes
        self._setup()
        hip.check(self._lib.xen_bits_expand_rows(
            bits, plane, colors, self.width, self.height, self.pitch,
            (ctypes.c_int32 * 3)(*dead), (ctypes.c_int32 * 3)(*alive), None),
            "xen_bits_expand_rows")

    def launch(self, src, dst, steps, born, sustain):
        """``steps`` (at most 8) steps from ``src`` into ``dst``."""
        self._setup()
        if self._rule != (born, sustain):
            # the table is read by launches already queued on the same
            # (null) stream; the copy is ordered after them
            self._table[0, 4:6] = (born, sustain)
            self._table_buf.upload(self._table)
            self._rule = (born, sustain)
        hip.check(self._lib.xen_life_batch_launch(
            src, dst, self.nbytes, self._table_buf.ptr, self._lanes.ptr,
            self._rows.ptr, self.n_waves, steps, None),
            "xen_life_batch_launch")
