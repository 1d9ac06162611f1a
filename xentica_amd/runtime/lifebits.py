"""
Bit-packed Life-like fields on the GPU (the fast engine of 2-D Moore tori).

A :class:`LifeBitsField` holds a ``width x height`` torus with one bit per
cell (32 cells per 32-bit word, x0 fastest -- the same bytes as 64 cells
per little-endian uint64) in two ping-pong buffers, and advances it with
``xen_life_bits_run``: ``depth`` steps per launch, kept in registers.  The
rule is the reference's Life-like rule (``examples/game_of_life.py:55-70,
190-197``): ``born = (B >> count) & 1``, ``sustained = (S >> count) & 1``,
``next = born | sustained & state``.
"""
import numpy as np

from xentica_amd.core.exceptions import XenticaException
from xentica_amd.runtime import checkpoint, hip, kernels

__all__ = ['LifeBitsField', 'Life3DField', 'Life3DBytes', ]


class LifeBitsField:
    """
    A packed Life-like torus resident in HBM.

    :param width: cells per row (a multiple of 32)
    :param height: rows
    :param born: birth mask (bit ``c`` set: born with ``c`` neighbours)
    :param sustain: survival mask
    :param depth: steps per launch (1..8)
    :param strips: row strips per launch (0 = automatic)
    """

    def __init__(self, width, height, born=8, sustain=12, depth=4, strips=0,
                 stream=None):
        if width % 32 or width <= 0 or height <= 0:
            raise XenticaException(
                "bit-packed fields need a width that is a positive multiple "
                "of 32 (got %dx%d)" % (width, height))
        self.width, self.height = int(width), int(height)
        self.words = self.width // 32
        self.n_cells = self.width * self.height
        self.born, self.sustain = int(born), int(sustain)
        self.depth, self.strips = int(depth), int(strips)
        self.stream = stream
        nbytes = self.n_cells // 8
        self._bufs = [hip.Buffer(nbytes), hip.Buffer(nbytes)]
        self._cur = 0

    @property
    def state_ptr(self):
        """Device address of the buffer holding the current state."""
        return self._bufs[self._cur].ptr

    @property
    def nbytes(self):
        """Bytes of one packed state."""
        return self.n_cells // 8

    def set_cells(self, cells):
        """Load a uint8 plane (one 0/1 byte per cell, x0 fastest)."""
        cells = np.ascontiguousarray(cells, dtype=np.uint8).reshape(-1)
        if cells.size != self.n_cells:
            raise XenticaException("expected %d cells, got %d" % (
                self.n_cells, cells.size))
        staging = hip.Buffer(self.n_cells)
        staging.upload(cells & 1, self.stream)
        kernels.bits_pack(staging.ptr, self.state_ptr, self.n_cells,
                          self.stream)
        self.synchronize()

    def set_bits(self, bits):
        """Load a packed state (``n_cells / 32`` uint32 words)."""
        bits = np.ascontiguousarray(bits, dtype=np.uint32).reshape(-1)
        self._bufs[self._cur].upload(bits, self.stream)

    def get_cells(self):
        """The state as a uint8 plane."""
        staging = hip.Buffer(self.n_cells)
        kernels.bits_unpack(self.state_ptr, staging.ptr, self.n_cells,
                            self.stream)
        out = np.empty(self.n_cells, dtype=np.uint8)
        staging.download(out, self.stream)
        return out

    def get_bits(self):
        """The packed state as uint32 words."""
        out = np.empty(self.n_cells // 32, dtype=np.uint32)
        self._bufs[self._cur].download(out, self.stream)
        return out

    def crc32(self):
        """``binascii.crc32`` of the uint8 plane, computed on the device."""
        return kernels.crc32_bits(self.state_ptr, self.n_cells, self.stream)

    def launch(self, depth=None):
        """One launch of ``depth`` steps (asynchronous)."""
        depth = self.depth if depth is None else depth
        kernels.life_bits_launch(
            self.state_ptr, self._bufs[1 - self._cur].ptr, self.words,
            self.height, self.born, self.sustain, depth, self.strips,
            self.stream)
        self._cur = 1 - self._cur

    def run(self, steps):
        """Advance ``steps`` steps (asynchronous)."""
        if steps <= 0:
            return
        in_b = kernels.life_bits_run(
            self.state_ptr, self._bufs[1 - self._cur].ptr, self.words,
            self.height, self.born, self.sustain, steps, self.depth,
            self.strips, self.stream)
        if in_b:
            self._cur = 1 - self._cur

    def save(self, filename):
        """Write the state in the reference's checkpoint format: plane 0 of
        a uint8 cells array with the 2-D Moore buffer planes (zero), zero
        colours and a fresh random stream (``xentica/core/base.py:755-763``);
        a ``GameOfLife`` of the same size loads it."""
        checkpoint.write(filename, checkpoint.field_state(
            self.get_cells(), checkpoint.MOORE_BUFFERS[2]))

    def load(self, filename):
        """Read plane 0 of a checkpoint (this field's or a model's)."""
        self.set_cells(checkpoint.main_plane(checkpoint.read(filename),
                                             self.n_cells))

    def synchronize(self):
        """Wait for queued work."""
        if self.stream is not None:
            self.stream.synchronize()
        else:
            hip.device().synchronize()


class Life3DField:
    """
    A packed 3-D Life-like torus with the reference's 3-D Moore list.

    The count of a cell is its 3x3x3 cube minus the cell at (x+1, y+1,
    z+1): the reference's ``MooreNeighborhood`` in 3-D keeps the centre and
    drops that offset (``xentica/core/topology/neighborhood.py:159-167``).
    32 cells per 32-bit word, x fastest; ``width / 32`` must divide 64.
    """

    def __init__(self, size, born, sustain, z_tiles=0, stream=None):
        width, height, depth = (int(v) for v in size)
        if width % 32 or 64 % (width // 32):
            raise XenticaException("width / 32 must divide 64")
        self.size = (width, height, depth)
        self.n_cells = width * height * depth
        self.born, self.sustain = int(born), int(sustain)
        self.z_tiles = int(z_tiles)
        self.stream = stream
        self._bufs = [hip.Buffer(self.n_cells // 8), hip.Buffer(self.n_cells // 8)]
        self._cur = 0

    @property
    def state_ptr(self):
        """Device address of the current state."""
        return self._bufs[self._cur].ptr

    def set_cells(self, cells, chunk=1 << 28):
        """Load a uint8 plane (x fastest), staged in chunks."""
        cells = np.ascontiguousarray(cells, dtype=np.uint8).reshape(-1)
        if cells.size != self.n_cells:
            raise XenticaException("expected %d cells" % self.n_cells)
        staging = hip.Buffer(min(chunk, self.n_cells))
        for start in range(0, self.n_cells, chunk):
            part = cells[start:start + chunk] & 1
            staging.upload(part, self.stream)
            kernels.bits_pack(staging.ptr, self.state_ptr + start // 8,
                              part.size, self.stream)
            self.synchronize()

    def get_cells(self):
        """The state as a uint8 plane."""
        staging = hip.Buffer(self.n_cells)
        kernels.bits_unpack(self.state_ptr, staging.ptr, self.n_cells,
                            self.stream)
        out = np.empty(self.n_cells, dtype=np.uint8)
        staging.download(out, self.stream)
        return out

    def crc32(self):
        """``binascii.crc32`` of the uint8 plane, on the device."""
        return kernels.crc32_bits(self.state_ptr, self.n_cells, self.stream)

    def step(self, steps=1):
        """Advance ``steps`` steps (asynchronous)."""
        for _ in range(steps):
            kernels.life3d_bits_step(
                self.state_ptr, self._bufs[1 - self._cur].ptr, self.size,
                self.born, self.sustain, self.z_tiles, self.stream)
            self._cur = 1 - self._cur

    def save(self, filename):
        """Write the state in the reference's checkpoint format (plane 0
        of a uint8 cells array with the 26 buffer planes of the 3-D Moore
        list, zero; zero colours; a fresh random stream)."""
        checkpoint.write(filename, checkpoint.field_state(
            self.get_cells(), checkpoint.MOORE_BUFFERS[3]))

    def load(self, filename):
        """Read plane 0 of a checkpoint (this field's or a model's)."""
        self.set_cells(checkpoint.main_plane(checkpoint.read(filename),
                                             self.n_cells))

    def synchronize(self):
        """Wait for queued work."""
        if self.stream is not None:
            self.stream.synchronize()
        else:
            hip.device().synchronize()


class Life3DBytes:
    """
    The same 3-D rule on one uint8 per cell (the reference's layout of a
    1-bit state, ``xentica/core/base.py:659-661``).

    ``direct=False``: the plane is packed once before the first step, the
    steps run on the bit-sliced kernel (widths as for
    :class:`Life3DField`), and the plane is expanded again when it is
    read (``state_ptr``, ``get_cells``, ``crc32``): one uint8 read and one
    write per cell for any number of steps.  ``direct=True``: one thread
    per cell reading its 26 neighbours, any width (an independent check).
    """

    def __init__(self, size, born, sustain, stream=None, direct=False):
        self.direct = bool(direct)
        self.size = tuple(int(v) for v in size)
        self.n_cells = self.size[0] * self.size[1] * self.size[2]
        self.born, self.sustain = int(born), int(sustain)
        self.stream = stream
        self._bufs = [hip.Buffer(self.n_cells), hip.Buffer(self.n_cells)]
        self._cur = 0
        self._packed = None if self.direct else Life3DField(
            self.size, born, sustain, stream=stream)
        self._in_bits = False

    def _plane(self):
        """The uint8 plane, expanded from the packed state if it is behind."""
        if self._in_bits:
            kernels.bits_unpack(self._packed.state_ptr,
                                self._bufs[self._cur].ptr, self.n_cells,
                                self.stream)
            self._in_bits = False
        return self._bufs[self._cur].ptr

    @property
    def state_ptr(self):
        """Device address of the current uint8 plane."""
        return self._plane()

    def set_cells(self, cells):
        """Load a uint8 plane."""
        self._in_bits = False
        self._bufs[self._cur].upload(
            np.ascontiguousarray(cells, dtype=np.uint8).reshape(-1) & 1,
            self.stream)

    def get_cells(self):
        """The state as a uint8 plane."""
        out = np.empty(self.n_cells, dtype=np.uint8)
        self._plane()
        self._bufs[self._cur].download(out, self.stream)
        return out

    def crc32(self):
        """``binascii.crc32`` of the plane, on the device."""
        return kernels.crc32_u8(self._plane(), self.n_cells, self.stream)

    def save(self, filename):
        """Write the state in the reference's checkpoint format (as
        :meth:`Life3DField.save`)."""
        checkpoint.write(filename, checkpoint.field_state(
            self.get_cells(), checkpoint.MOORE_BUFFERS[3]))

    def load(self, filename):
        """Read plane 0 of a checkpoint (this field's or a model's)."""
        self.set_cells(checkpoint.main_plane(checkpoint.read(filename),
                                             self.n_cells))

    def step(self, steps=1):
        """Advance ``steps`` steps (asynchronous)."""
        if steps <= 0:
            return
        if self.direct:
            for _ in range(steps):
                kernels.life3d_u8_direct_step(
                    self.state_ptr, self._bufs[1 - self._cur].ptr, self.size,
                    self.born, self.sustain, self.stream)
                self._cur = 1 - self._cur
            return
        if not self._in_bits:
            kernels.bits_pack(self._bufs[self._cur].ptr,
                              self._packed.state_ptr, self.n_cells,
                              self.stream)
            self._in_bits = True
        self._packed.step(steps)

    def synchronize(self):
        """Wait for queued work."""
        if self.stream is not None:
            self.stream.synchronize()
        else:
            hip.device().synchronize()
