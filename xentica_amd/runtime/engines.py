"""
Step engines: how a model's step is executed on the GPU.

* ``plain`` -- the reference's structure (``xentica/core/base.py:703-728``):
  one ``emit`` launch writing every buffer plane to HBM, then one
  ``absorb`` launch reading them, per step.  Works for every model the DSL
  can express and is the parity bridge for the faster engines.
* ``fused`` -- emit and absorb fused in one generated kernel with the
  buffers in LDS and several steps per launch (``runtime/fused.py``);
  steps are deferred until a launch is full or the state is read.
* ``fused_unfolded`` -- ``fused`` with buffer reads kept on LDS buffer
  planes even when emit only copies the state (the general form of the
  kernel; kept selectable to check the folded one against it).
* ``fused_generic`` -- ``fused`` with every cell of a step going through
  one generic loop over the valid rectangle (the form used for tiles whose
  rows do not split evenly over the block; kept selectable to check the
  split form against it).
* ``bits`` -- the bit-sliced Life-like kernels (``runtime/bits_engine.py``)
  for models whose generated code is, exactly, a Life-like rule on a
  1-bit state on a 2-D or 3-D Moore torus.
* ``auto`` -- ``bits`` when the model qualifies, else ``fused`` when it
  qualifies, else ``plain``.

Engines share one interface: ``step()`` (advance one step, possibly
deferred), ``flush()`` (finish deferred steps), ``load_state(cells,
colors)`` and ``launch_elementwise(kernel, args, n)``.
"""
import numpy as np

from xentica_amd.core.exceptions import XenticaException

__all__ = ['create', 'PlainEngine', 'ENGINES', 'elementwise_grid']

#: Threads per workgroup of the grid-stride kernels (4 wave64s).
BLOCK = 256
#: Upper bound on workgroups of a grid-stride launch: 64 per CU on 256 CUs.
MAX_GRID = 256 * 64


def elementwise_grid(count):
    """Workgroup count of a grid-stride launch over ``count`` items."""
    return max(1, min((count + BLOCK - 1) // BLOCK, MAX_GRID))


class PlainEngine:
    """Reference-structured emit/absorb launches with buffers in HBM."""

    name = "plain"

    def __init__(self, model):
        self.model = model
        if model.cells_num * (len(model.buffers) + 1) >= 2 ** 62:
            raise XenticaException("field too large for the plain engine")
        if model.cells_num >= 2 ** 32:
            raise XenticaException(
                "the plain engine indexes cells with 32-bit unsigned ints")
        self._key = None
        self._emit = self._absorb = None

    def launch_elementwise(self, kernel, args, count):
        """Launch a grid-stride kernel over ``count`` items."""
        kernel(*args, grid=(elementwise_grid(count), 1, 1),
               block=(BLOCK, 1, 1))

    def step(self):
        """Launch ``emit`` then ``absorb`` with current parameter values.

        The two launches are prepared once (arguments converted, pointer
        arrays built) and reused while the parameter values and the device
        arrays stay the same: at small fields the per-step host work is
        most of a step.
        """
        model = self.model
        arrays = model.gpu.arrays
        key = (tuple(param.value for param in model.emit_params),
               tuple(param.value for param in model.absorb_params),
               arrays.cells.device_ptr(), arrays.colors.device_ptr())
        if key != self._key:
            kernels = model.gpu.kernels
            cells_num = np.int64(model.cells_num)
            grid = (elementwise_grid(model.cells_num), 1, 1)
            args = [param.dtype(param.value) for param in model.emit_params]
            self._emit = kernels.emit.prepare(
                arrays.cells, *args, cells_num, grid=grid, block=(BLOCK, 1, 1))
            args = [param.dtype(param.value)
                    for param in model.absorb_params]
            self._absorb = kernels.absorb.prepare(
                arrays.cells, arrays.colors, *args, cells_num, grid=grid,
                block=(BLOCK, 1, 1))
            self._key = key
        self._emit()
        self._absorb()
        # the colour bounds the fused engine relies on are not followed here
        arrays.colors.bounds = None

    def flush(self):
        """Nothing is deferred: launches are already queued in order."""

    def load_state(self, cells, colors):
        """Replace the device arrays' contents."""
        arrays = self.model.gpu.arrays
        arrays.cells.set(cells)
        arrays.colors.set(colors)


def _fused(model):
    from xentica_amd.runtime.fused import FusedEngine
    return FusedEngine(model)


def _fused_unfolded(model):
    from xentica_amd.runtime.fused import FusedEngine
    return FusedEngine(model, fold=False)


def _fused_generic(model):
    from xentica_amd.runtime.fused import FusedEngine
    return FusedEngine(model, split=False)


def _bits(model):
    from xentica_amd.runtime.bits_engine import BitsEngine
    return BitsEngine(model)


def _fused_or_plain(model):
    """The fused engine when the model qualifies, else the plain one."""
    from xentica_amd.runtime.fused import FusedEngine, eligible
    ok, _ = eligible(model)
    if ok:
        try:
            return FusedEngine(model)
        except XenticaException:
            pass
    return PlainEngine(model)


def _auto(model):
    """The bit-sliced engine, else the fused one, else the plain one."""
    from xentica_amd.runtime.bits_engine import BitsEngine, eligible_size
    if eligible_size(model.size):
        try:
            return BitsEngine(model)
        except XenticaException:
            pass
    return _fused_or_plain(model)


ENGINES = {
    'plain': PlainEngine,
    'fused': _fused,
    'fused_unfolded': _fused_unfolded,
    'fused_generic': _fused_generic,
    'bits': _bits,
    'fused_or_plain': _fused_or_plain,
    'auto': _auto,
}


def create(model, name):
    """Build the engine ``name`` for ``model``."""
    if name not in ENGINES:
        raise XenticaException("unknown engine %r (known: %s)" % (
            name, ", ".join(sorted(ENGINES))))
    return ENGINES[name](model)
