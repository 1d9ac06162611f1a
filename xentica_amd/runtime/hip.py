"""
ctypes binding of ``libxentica_hip.so`` (the C ABI in ``include/xentica_hip.h``).

This is the only module that talks to the HIP runtime.  It replaces
``pycuda.autoinit``, ``SourceModule`` and ``gpuarray`` of the reference
(``xentica/core/base.py:94-96, 527-569``).  There is no fallback: when the
library is missing or no GPU is visible every entry point raises
:class:`XenticaException`.

Compiled code objects are cached on disk, keyed by a hash of the source
and the compiler options, so a model class compiles once per machine.
"""
import ctypes
import gc
import hashlib
import os
import re
import threading

import numpy as np

from xentica_amd.core.exceptions import XenticaException

__all__ = [
    'library', 'check', 'Device', 'device', 'Module', 'Function',
    'Buffer', 'Stream', 'Event', 'Graph', 'compile_source',
    'COMPILE_OPTIONS', 'compile_options', 'library_path',
]

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB_NAME = "libxentica_hip.so"

#: hiprtc options of every generated kernel: gfx950, -O3, no packed FP32
#: instructions, and no contraction of a*b+c into FMA (C semantics).
COMPILE_OPTIONS = (
    "--offload-arch=gfx950", "-O3", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-w",
)


def compile_options(fp_contract="off"):
    """hiprtc options with the given FP contraction mode (off/on/fast)."""
    if fp_contract not in ("off", "on", "fast"):
        raise XenticaException("fp_contract must be off, on or fast")
    return tuple("-ffp-contract=%s" % fp_contract
                 if opt.startswith("-ffp-contract=") else opt
                 for opt in COMPILE_OPTIONS)

_LOCK = threading.RLock()
_LIB = None

_vp = ctypes.c_void_p
_SIGNATURES = {
    'xen_init': [ctypes.c_int],
    'xen_device_count': [ctypes.POINTER(ctypes.c_int)],
    'xen_device_name': [ctypes.c_char_p, ctypes.c_int],
    'xen_compile': [ctypes.c_char_p, ctypes.c_char_p,
                    ctypes.POINTER(ctypes.c_char_p), ctypes.c_int,
                    ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_size_t)],
    'xen_module_load': [_vp, ctypes.POINTER(_vp)],
    'xen_module_unload': [_vp],
    'xen_module_function': [_vp, ctypes.c_char_p, ctypes.POINTER(_vp)],
    'xen_launch': [_vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                   ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                   ctypes.c_uint, _vp, ctypes.POINTER(_vp)],
    'xen_occupancy': [_vp, ctypes.c_uint, ctypes.c_uint,
                      ctypes.POINTER(ctypes.c_int),
                      ctypes.POINTER(ctypes.c_int)],
    'xen_malloc': [ctypes.POINTER(_vp), ctypes.c_size_t],
    'xen_free': [_vp],
    'xen_memcpy_h2d': [_vp, _vp, ctypes.c_size_t, _vp],
    'xen_memcpy_d2h': [_vp, _vp, ctypes.c_size_t, _vp],
    'xen_memcpy_d2d': [_vp, _vp, ctypes.c_size_t, _vp],
    'xen_memset': [_vp, ctypes.c_int, ctypes.c_size_t, _vp],
    'xen_mem_info': [ctypes.POINTER(ctypes.c_size_t),
                     ctypes.POINTER(ctypes.c_size_t)],
    'xen_stream_create': [ctypes.POINTER(_vp)],
    'xen_stream_destroy': [_vp],
    'xen_stream_sync': [_vp],
    'xen_device_sync': [],
    'xen_event_create': [ctypes.POINTER(_vp)],
    'xen_event_destroy': [_vp],
    'xen_event_record': [_vp, _vp],
    'xen_event_sync': [_vp],
    'xen_event_elapsed': [_vp, _vp, ctypes.POINTER(ctypes.c_float)],
    'xen_graph_begin': [_vp],
    'xen_graph_end': [_vp, ctypes.POINTER(_vp)],
    'xen_graph_launch': [_vp, _vp],
    'xen_graph_destroy': [_vp],
}


def library_path():
    """Path of the in-tree shared library."""
    return os.path.join(_PKG_DIR, _LIB_NAME)


def _share_torch_runtime():
    """
    Load PyTorch's HIP runtime before this library, when PyTorch is there.

    PyTorch ships its own ``libamdhip64.so`` under the same soname as the
    system one; the dynamic loader keeps whichever comes first for the
    whole process.  Loading PyTorch first gives one HIP runtime that both
    its tensors (device memory, streams) and these kernels use; the other
    order leaves PyTorch unable to see the GPU.
    """
    try:
        import torch  # noqa: F401  pylint: disable=import-outside-toplevel
    except ImportError:
        pass


def library():
    """Load ``libxentica_hip.so`` once; raise if it was never built."""
    global _LIB  # pylint: disable=global-statement
    if _LIB is not None:
        return _LIB
    with _LOCK:
        if _LIB is None:
            path = library_path()
            if not os.path.exists(path):
                raise XenticaException(
                    "%s is not built: run __graft_entry__.build() "
                    "(hipcc --offload-arch=gfx950)" % path)
            _share_torch_runtime()
            lib = ctypes.CDLL(path)
            for name, argtypes in _SIGNATURES.items():
                func = getattr(lib, name)
                func.argtypes = argtypes
                func.restype = ctypes.c_int
            lib.xen_last_error.argtypes = []
            lib.xen_last_error.restype = ctypes.c_char_p
            lib.xen_code_free.argtypes = [_vp]
            lib.xen_code_free.restype = None
            _LIB = lib
    return _LIB


def check(status, what=""):
    """Raise :class:`XenticaException` with the library's message."""
    if status != 0:
        msg = library().xen_last_error().decode(errors="replace")
        raise XenticaException("%s%s (code %d)" % (
            (what + ": ") if what else "", msg, status))


class Device:
    """The one HIP device this process computes on."""

    def __init__(self, index=0):
        self.index = index
        lib = library()
        count = ctypes.c_int(0)
        check(lib.xen_device_count(ctypes.byref(count)), "no HIP device")
        if count.value < 1:
            raise XenticaException("no HIP device visible")
        check(lib.xen_init(index), "xen_init")
        buf = ctypes.create_string_buffer(256)
        check(lib.xen_device_name(buf, 256))
        self.name = buf.value.decode()
        # xen_device_name ends with "(<arch>, <n> CUs)"
        found = re.search(r"(\d+) CUs\)$", self.name)
        self.compute_units = int(found.group(1)) if found else 256

    def synchronize(self):
        """Wait for all work on the device."""
        check(library().xen_device_sync())

    def mem_info(self):
        """``(free_bytes, total_bytes)`` of device memory."""
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        check(library().xen_mem_info(ctypes.byref(free), ctypes.byref(total)))
        return free.value, total.value


_DEVICE = None


def device():
    """The process-wide :class:`Device` (initialised on first use)."""
    global _DEVICE  # pylint: disable=global-statement
    if _DEVICE is None:
        with _LOCK:
            if _DEVICE is None:
                _DEVICE = Device(int(os.environ.get("XENTICA_DEVICE", "0")))
    return _DEVICE


def _cache_dir():
    root = os.environ.get("XENTICA_CACHE_DIR") or os.path.join(
        os.path.expanduser("~"), ".cache", "xentica_amd")
    os.makedirs(root, exist_ok=True)
    return root


def compile_source(source, name="xentica.hip", options=COMPILE_OPTIONS):
    """Compile HIP source for gfx950 with hiprtc; returns the code object."""
    key = hashlib.sha256(
        ("\0".join(options) + "\0" + source).encode()).hexdigest()[:32]
    path = os.path.join(_cache_dir(), key + ".co")
    if os.path.exists(path):
        with open(path, "rb") as cached:
            return cached.read()
    lib = library()
    opts = (ctypes.c_char_p * len(options))(*[o.encode() for o in options])
    code, size = _vp(), ctypes.c_size_t()
    check(lib.xen_compile(source.encode(), name.encode(), opts, len(options),
                          ctypes.byref(code), ctypes.byref(size)),
          "compiling %s" % name)
    try:
        blob = ctypes.string_at(code, size.value)
    finally:
        lib.xen_code_free(code)
    tmp = "%s.%d.tmp" % (path, os.getpid())
    with open(tmp, "wb") as out:
        out.write(blob)
    os.replace(tmp, path)
    return blob


def _compile_into_cache(source, options):
    """Worker of :func:`precompile`: compile one source into the cache."""
    compile_source(source, "precompiled.hip", tuple(options))
    return True


def precompile(sources, options=COMPILE_OPTIONS, workers=0):
    """
    Compile many sources into the on-disk code-object cache at once.

    hiprtc serialises compilations within one process, so the sources go
    to worker processes (``workers``, default: the usable cores, at most
    16); the workers only compile and never touch the GPU.  Sources
    already cached are skipped.  Later :func:`compile_source` calls of
    the same source and options read the cache.
    """
    options = tuple(options)
    todo = []
    seen = set()
    for source in sources:
        key = hashlib.sha256(
            ("\0".join(options) + "\0" + source).encode()).hexdigest()[:32]
        if key in seen or os.path.exists(
                os.path.join(_cache_dir(), key + ".co")):
            continue
        seen.add(key)
        todo.append(source)
    if not todo:
        return 0
    if len(todo) == 1:
        compile_source(todo[0], "precompiled.hip", options)
        return 1
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor
    workers = workers or min(16, len(os.sched_getaffinity(0)))
    workers = max(1, min(workers, len(todo)))
    with ProcessPoolExecutor(
            workers, mp_context=multiprocessing.get_context("spawn")) as pool:
        list(pool.map(_compile_into_cache, todo, [options] * len(todo),
                      chunksize=max(1, len(todo) // (4 * workers))))
    return len(todo)


def _as_ctype(arg):
    """Convert one kernel argument to a ctypes object."""
    if hasattr(arg, "device_ptr"):
        return _vp(arg.device_ptr())
    if isinstance(arg, (bool, np.bool_)):
        return ctypes.c_bool(bool(arg))
    if isinstance(arg, np.generic):
        kind, size = arg.dtype.kind, arg.dtype.itemsize
        if kind == 'f':
            return ctypes.c_float(arg) if size == 4 else ctypes.c_double(arg)
        table = {
            ('i', 4): ctypes.c_int32, ('i', 8): ctypes.c_int64,
            ('u', 4): ctypes.c_uint32, ('u', 8): ctypes.c_uint64,
            ('i', 2): ctypes.c_int16, ('u', 2): ctypes.c_uint16,
            ('i', 1): ctypes.c_int8, ('u', 1): ctypes.c_uint8,
        }
        return table[(kind, size)](int(arg))
    if isinstance(arg, ctypes._SimpleCData):  # pylint: disable=protected-access
        return arg
    if isinstance(arg, int):
        return ctypes.c_int32(arg)
    if isinstance(arg, float):
        return ctypes.c_float(arg)
    raise XenticaException("unsupported kernel argument %r" % (arg,))


class Function:
    """A kernel of a loaded module."""

    def __init__(self, module, name, handle):
        self.module = module
        self.name = name
        self.handle = handle

    def __call__(self, *args, grid=(1, 1, 1), block=(256, 1, 1), shared=0,
                 stream=None):
        """Launch with positional arguments (pointers and NumPy scalars)."""
        cargs = [_as_ctype(arg) for arg in args]
        ptrs = (_vp * max(1, len(cargs)))(
            *[ctypes.cast(ctypes.pointer(c), _vp) for c in cargs])
        grid = tuple(grid) + (1,) * (3 - len(grid))
        block = tuple(block) + (1,) * (3 - len(block))
        check(library().xen_launch(
            self.handle, grid[0], grid[1], grid[2], block[0], block[1],
            block[2], shared, stream.handle if stream else None, ptrs),
            "launching %s" % self.name)


    def occupancy(self, block=256, shared=0):
        """``(workgroups one CU holds at once, CUs)`` for ``block``-thread
        workgroups of this kernel (registers and LDS counted)."""
        # kept by the module (functions are made afresh by each lookup)
        cache = self.module.__dict__.setdefault("_occupancy", {})
        key = (self.name, block, shared)
        if key not in cache:
            per_cu, cus = ctypes.c_int(), ctypes.c_int()
            check(library().xen_occupancy(
                self.handle, block, shared, ctypes.byref(per_cu),
                ctypes.byref(cus)), "occupancy of %s" % self.name)
            cache[key] = (per_cu.value, cus.value)
        return cache[key]

    def prepare(self, *args, grid=(1, 1, 1), block=(256, 1, 1), shared=0,
                stream=None):
        """A launch whose arguments are converted once; call it to launch
        (for launches repeated with the same arguments, e.g. every step)."""
        return PreparedLaunch(self, args, grid, block, shared, stream)


class PreparedLaunch:
    """A kernel launch with its argument array built in advance."""

    def __init__(self, function, args, grid, block, shared, stream):
        self.function = function  # keeps the module loaded
        self._cargs = [_as_ctype(arg) for arg in args]
        self._ptrs = (_vp * max(1, len(self._cargs)))(
            *[ctypes.cast(ctypes.pointer(c), _vp) for c in self._cargs])
        grid = tuple(grid) + (1,) * (3 - len(grid))
        block = tuple(block) + (1,) * (3 - len(block))
        self._stream = stream  # keeps the stream alive
        self._call = (function.handle, grid[0], grid[1], grid[2], block[0],
                      block[1], block[2], shared,
                      stream.handle if stream else None, self._ptrs)
        self._what = "launching %s" % function.name

    def __call__(self):
        status = library().xen_launch(*self._call)
        if status:
            check(status, self._what)


class Module:
    """A loaded code object."""

    def __init__(self, code):
        device()
        self._code = code
        handle = _vp()
        check(library().xen_module_load(code, ctypes.byref(handle)),
              "loading module")
        self.handle = handle
        # kernel handles only: a Function keeps its Module alive, and a
        # Module that held its Functions would form a reference cycle whose
        # unload waits for the cyclic collector (which may run while a
        # stream is being captured, where an unload is not allowed)
        self._handles = {}

    @classmethod
    def from_source(cls, source, name="xentica.hip"):
        """Compile ``source`` with hiprtc and load it."""
        return cls(compile_source(source, name))

    def get_function(self, name):
        """Look up a kernel by its (``extern "C"``) name."""
        if name not in self._handles:
            handle = _vp()
            check(library().xen_module_function(
                self.handle, name.encode(), ctypes.byref(handle)))
            self._handles[name] = handle
        return Function(self, name, self._handles[name])

    def __del__(self):
        if getattr(self, "handle", None) and _LIB is not None:
            _LIB.xen_module_unload(self.handle)


class Buffer:
    """A device allocation, freed with the object."""

    def __init__(self, nbytes):
        device()
        self.nbytes = int(nbytes)
        ptr = _vp()
        check(library().xen_malloc(ctypes.byref(ptr), self.nbytes),
              "allocating %d bytes" % self.nbytes)
        self.ptr = ptr.value

    def device_ptr(self):
        """Integer device address."""
        return self.ptr

    def upload(self, host, stream=None):
        """Copy a contiguous NumPy array into the buffer (synchronous)."""
        host = np.ascontiguousarray(host)
        if host.nbytes > self.nbytes:
            raise XenticaException("upload of %d bytes into %d-byte buffer" % (
                host.nbytes, self.nbytes))
        check(library().xen_memcpy_h2d(
            self.ptr, host.ctypes.data, host.nbytes,
            stream.handle if stream else None))

    def download(self, host, stream=None):
        """Copy the buffer's first ``host.nbytes`` bytes into ``host``."""
        if host.nbytes > self.nbytes or not host.flags.c_contiguous:
            raise XenticaException("bad download target")
        check(library().xen_memcpy_d2h(
            host.ctypes.data, self.ptr, host.nbytes,
            stream.handle if stream else None))

    def fill_zero(self, stream=None):
        """Zero the whole buffer."""
        check(library().xen_memset(self.ptr, 0, self.nbytes,
                                   stream.handle if stream else None))

    def __del__(self):
        if getattr(self, "ptr", None) and _LIB is not None:
            _LIB.xen_free(self.ptr)
            self.ptr = None


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self):
        device()
        handle = _vp()
        check(library().xen_stream_create(ctypes.byref(handle)))
        self.handle = handle.value

    def synchronize(self):
        """Wait for the stream's work."""
        check(library().xen_stream_sync(self.handle))

    def __del__(self):
        if getattr(self, "handle", None) and _LIB is not None:
            _LIB.xen_stream_destroy(self.handle)


class Event:
    """A HIP event for timing on a stream."""

    def __init__(self):
        device()
        handle = _vp()
        check(library().xen_event_create(ctypes.byref(handle)))
        self.handle = handle.value

    def record(self, stream=None):
        """Record the event on ``stream``."""
        check(library().xen_event_record(
            self.handle, stream.handle if stream else None))

    def synchronize(self):
        """Wait for the event."""
        check(library().xen_event_sync(self.handle))

    def elapsed_ms(self, stop):
        """Milliseconds from this event to ``stop``."""
        val = ctypes.c_float()
        check(library().xen_event_elapsed(self.handle, stop.handle,
                                          ctypes.byref(val)))
        return val.value

    def __del__(self):
        if getattr(self, "handle", None) and _LIB is not None:
            _LIB.xen_event_destroy(self.handle)


class Graph:
    """A captured and instantiated HIP graph."""

    def __init__(self, handle):
        self.handle = handle

    _gc_was_enabled = None

    @classmethod
    def begin(cls, stream):
        """
        Start capturing launches on ``stream``.

        Python's cyclic garbage collector is paused until :meth:`end`: an
        object it frees runs ``__del__`` (``hipFree``, module unload),
        which is not allowed on a thread that is capturing.
        """
        cls._gc_was_enabled = gc.isenabled()
        gc.disable()
        try:
            check(library().xen_graph_begin(stream.handle), "graph capture")
        except BaseException:
            cls._restore_gc()
            raise

    @classmethod
    def _restore_gc(cls):
        if cls._gc_was_enabled:
            gc.enable()
        cls._gc_was_enabled = None

    @classmethod
    def end(cls, stream):
        """Finish capturing on ``stream`` and instantiate the graph."""
        handle = _vp()
        try:
            check(library().xen_graph_end(stream.handle,
                                          ctypes.byref(handle)),
                  "graph instantiate")
        finally:
            cls._restore_gc()
        return cls(handle.value)

    def launch(self, stream):
        """Replay the graph on ``stream``."""
        check(library().xen_graph_launch(self.handle, stream.handle))

    def __del__(self):
        if getattr(self, "handle", None) and _LIB is not None:
            _LIB.xen_graph_destroy(self.handle)
