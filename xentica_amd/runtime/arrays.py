"""
Device arrays with the part of PyCUDA's ``GPUArray`` surface the reference
uses: ``.get()`` returning a NumPy copy, ``.shape``, ``.dtype``, ``.size``
(``xentica/core/base.py:544-569, 753, 759-760``; every reference test
reads ``model.gpu.arrays.cells.get()``).

An array stays on the host until a kernel first needs its device address,
so models can be built, seeded and inspected on machines without a GPU.
An owner may attach a ``before_read`` hook that runs ahead of every
``.get()`` and ``.set()``; the engines use it to finish deferred steps
first, so that a read sees them and a write is not overwritten by them.

An array of interleaved channels (the colours: red, green, blue per cell)
can keep per-channel value bounds: computed from the data at every
``set()``, updated by the engines after their kernels write, ``None``
when unknown.  Code that writes the array through its device address must
set ``bounds`` to ``None``.
"""
import numpy as np

from xentica_amd.runtime import hip

__all__ = ['DeviceArray', 'to_gpu', 'zeros', ]


class DeviceArray:
    """A flat typed array living in HBM."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.shape = host.shape
        self.dtype = host.dtype
        self._host = host.copy()
        self._buffer = None
        self.before_read = None
        self.channels = None
        self.bounds = None

    def track_bounds(self, channels):
        """Keep per-channel ``(min, max)`` bounds of the contents."""
        self.channels = channels
        self.bounds = (None if self._host is None
                       else _channel_bounds(self._host, channels))

    @property
    def size(self):
        """Number of elements."""
        return int(np.prod(self.shape, dtype=np.int64))

    @property
    def nbytes(self):
        """Size in bytes."""
        return self.size * self.dtype.itemsize

    @property
    def on_device(self):
        """``True`` once the data lives in device memory."""
        return self._buffer is not None

    def device_ptr(self):
        """Device address, uploading the host copy on first use."""
        if self._buffer is None:
            buf = hip.Buffer(self.nbytes)
            buf.upload(self._host)
            self._buffer = buf
            self._host = None
        return self._buffer.ptr

    def get(self):
        """Return a NumPy copy of the contents."""
        if self.before_read is not None:
            self.before_read()
        if self._buffer is None:
            return self._host.copy()
        out = np.empty(self.shape, dtype=self.dtype)
        self._buffer.download(out)
        return out

    def set(self, host):
        """Replace the contents with ``host`` (same size and dtype)."""
        host = np.ascontiguousarray(host, dtype=self.dtype).reshape(self.shape)
        if self.before_read is not None:
            self.before_read()
        if self._buffer is None:
            self._host = host.copy()
        else:
            self._buffer.upload(host)
        if self.channels:
            self.bounds = _channel_bounds(host, self.channels)

    def __repr__(self):
        where = "device" if self.on_device else "host"
        return "DeviceArray(shape=%s, dtype=%s, %s)" % (
            self.shape, self.dtype, where)


def _channel_bounds(host, channels):
    """``((min, max), ...)`` of each channel of interleaved ``host``."""
    view = host.reshape(-1, channels)
    if not len(view):
        return None
    return tuple((int(view[:, c].min()), int(view[:, c].max()))
                 for c in range(channels))


def to_gpu(host):
    """Make a :class:`DeviceArray` holding a copy of ``host``."""
    return DeviceArray(host)


def zeros(shape, dtype):
    """Make a zero-filled :class:`DeviceArray`."""
    return DeviceArray(np.zeros(shape, dtype=dtype))
