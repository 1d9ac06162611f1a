"""
Device arrays with the part of PyCUDA's ``GPUArray`` surface the reference
uses: ``.get()`` returning a NumPy copy, ``.shape``, ``.dtype``, ``.size``
(``xentica/core/base.py:544-569, 753, 759-760``; every reference test
reads ``model.gpu.arrays.cells.get()``).

An array stays on the host until a kernel first needs its device address,
so models can be built, seeded and inspected on machines without a GPU.

An array may hold only a leading part of its contents, the rest being
zero: a model's cells array is ``n * (k + 1) + 1`` words of which only the
main-state plane (the first ``n``) is written by seeds; the buffer planes
start as zeros and only the plain engine ever uses them.  ``device_ptr()``
gives the whole array on the device (allocating and zeroing the tail),
``plane_ptr(count)`` a device copy of the first ``count`` elements only,
which is all the fused and bit-sliced engines touch: a 1024^3 3-D model
then needs 1 GiB of device memory for its state instead of 27 GiB.

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
    """
    A flat typed array living in HBM.

    :param host: the contents, or their leading part when ``length`` is
        given (the remaining elements are zero)
    :param length: the number of elements, if more than ``len(host)``
    """

    def __init__(self, host, length=None):
        host = np.ascontiguousarray(host)
        if length is None:
            self.shape = host.shape
        else:
            host = host.reshape(-1)
            if length < host.size:
                raise ValueError("length shorter than the given contents")
            self.shape = (int(length), )
        self.dtype = host.dtype
        self._host = host.reshape(-1).copy()   # leading part; rest zero
        self._buffer = None                     # the whole array
        self._plane = None                      # (buffer, count) prefix
        self.before_read = None
        self.channels = None
        self.bounds = None

    def track_bounds(self, channels):
        """Keep per-channel ``(min, max)`` bounds of the contents."""
        self.channels = channels
        if self._host is None:
            self.bounds = None
        elif not self._host.size:
            self.bounds = ((0, 0), ) * channels if self.size else None
        else:
            self.bounds = _channel_bounds(self._padded_host(), channels)

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
        """``True`` once (part of) the data lives in device memory."""
        return self._buffer is not None or self._plane is not None

    def _padded_host(self):
        if self._host.size == self.size:
            return self._host.reshape(self.shape)
        out = np.zeros(self.size, dtype=self.dtype)
        out[:self._host.size] = self._host
        return out.reshape(self.shape)

    def _fill(self, ptr, count):
        """Write the first ``count`` elements of the contents at ``ptr``
        from the host copy (which is then dropped once the whole array is
        on the device)."""
        item = self.dtype.itemsize
        known = min(count, self._host.size)
        if known:
            hip.check(hip.library().xen_memcpy_h2d(
                ptr, self._host.ctypes.data, known * item, None))
        if count > known:
            hip.check(hip.library().xen_memset(
                ptr + known * item, 0, (count - known) * item, None))
            # work on other (non-blocking) streams may use the address next
            hip.device().synchronize()

    def device_ptr(self):
        """Device address of the whole array (uploaded on first use)."""
        if self._buffer is None:
            buf = hip.Buffer(self.nbytes)
            if self._plane is not None:
                plane, count = self._plane
                item = self.dtype.itemsize
                hip.check(hip.library().xen_memcpy_d2d(
                    buf.ptr, plane.ptr, count * item, None))
                if self.size > count:
                    hip.check(hip.library().xen_memset(
                        buf.ptr + count * item, 0,
                        (self.size - count) * item, None))
                hip.device().synchronize()
                self._plane = None
            else:
                self._fill(buf.ptr, self.size)
            self._buffer = buf
            self._host = None
        return self._buffer.ptr

    def plane_ptr(self, count):
        """Device address of (at least) the first ``count`` elements; the
        rest of the array is not allocated on the device when it is not
        there yet.  Writes through this address must stay in the first
        ``count`` elements."""
        if self._buffer is not None:
            return self._buffer.ptr
        if self._plane is not None and self._plane[1] >= count:
            return self._plane[0].ptr
        if self._plane is not None or self._host.size > count:
            # the contents beyond the plane are not all zero, or a shorter
            # plane is there already: keep the whole array instead
            return self.device_ptr()
        buf = hip.Buffer(count * self.dtype.itemsize)
        self._fill(buf.ptr, count)
        self._plane = (buf, count)
        return buf.ptr

    def get(self):
        """Return a NumPy copy of the contents."""
        if self.before_read is not None:
            self.before_read()
        if self._buffer is not None:
            out = np.empty(self.shape, dtype=self.dtype)
            self._buffer.download(out)
            return out
        if self._plane is not None:
            plane, count = self._plane
            out = np.zeros(self.size, dtype=self.dtype)
            plane.download(out[:count])
            return out.reshape(self.shape)
        return self._padded_host().copy()

    def set(self, host):
        """Replace the contents with ``host`` (same size and dtype)."""
        host = np.ascontiguousarray(host, dtype=self.dtype).reshape(self.shape)
        if self.before_read is not None:
            self.before_read()
        flat = host.reshape(-1)
        if self._buffer is not None:
            self._buffer.upload(host)
        elif self._plane is not None and \
                not flat[self._plane[1]:].view(np.uint8).any():
            self._plane[0].upload(flat[:self._plane[1]])
        else:
            self._plane = None
            self._host = flat.copy()
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


def to_gpu(host, length=None):
    """Make a :class:`DeviceArray` holding a copy of ``host`` (followed by
    zeros up to ``length`` elements, when given)."""
    return DeviceArray(host, length)


def zeros(shape, dtype):
    """Make a zero-filled :class:`DeviceArray`."""
    count = int(np.prod(shape, dtype=np.int64))
    arr = DeviceArray(np.zeros(0, dtype=dtype), count)
    arr.shape = tuple(shape) if isinstance(shape, tuple) else (count, )
    return arr
