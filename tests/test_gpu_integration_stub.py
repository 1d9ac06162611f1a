"""
The ctypes binding shown to a maintainer of the reference (the stub that
replaces PyCUDA's SourceModule, to_gpu and launches), run as written
against this library: the generated source of a reference model compiled
with xen_compile, its emit/absorb launched through xen_launch, the state
compared with the plain engine after 10 steps.
"""
import ctypes

import numpy as np
import pytest

from examples.game_of_life import GameOfLife, GOLExperiment
from xentica_amd.runtime import hip

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def test_reference_side_stub_runs_generated_kernels():
    _lib = hip.library()  # libxentica_hip.so, loaded after PyTorch

    def _check(rc):
        if rc:
            raise RuntimeError(_lib.xen_last_error().decode())

    _check(_lib.xen_init(0))

    class SourceModule:
        OPTS = [b"--offload-arch=gfx950", b"-O3", b"-ffp-contract=off", b"-w"]

        def __init__(self, source):
            code, size = _vp(), ctypes.c_size_t()
            opts = (ctypes.c_char_p * len(self.OPTS))(*self.OPTS)
            _check(_lib.xen_compile(source.encode(), b"bsca.hip", opts,
                                    len(self.OPTS), ctypes.byref(code),
                                    ctypes.byref(size)))
            self.mod = _vp()
            try:  # the module keeps its own copy of the code object
                _check(_lib.xen_module_load(code, ctypes.byref(self.mod)))
            finally:
                _lib.xen_code_free(code)

        def get_function(self, name):
            fn = _vp()
            _check(_lib.xen_module_function(self.mod, name.encode(),
                                            ctypes.byref(fn)))
            return fn

    def to_gpu(host):
        ptr = _vp()
        _check(_lib.xen_malloc(ctypes.byref(ptr), ctypes.c_size_t(host.nbytes)))
        _check(_lib.xen_memcpy_h2d(ptr, _vp(host.ctypes.data),
                                   ctypes.c_size_t(host.nbytes), None))
        return ptr

    def launch(fn, args, n, block=256):
        cargs = [ctypes.c_void_p(a.value) if isinstance(a, _vp) else a
                 for a in args]
        ptrs = (_vp * len(cargs))(*[ctypes.cast(ctypes.pointer(c), _vp)
                                    for c in cargs])
        grid = (n + block - 1) // block
        _check(_lib.xen_launch(fn, grid, 1, 1, block, 1, 1, 0, None, ptrs))

    model = GameOfLife(type("StubGOL", (GOLExperiment, ), {"engine": "plain"}))
    n = model.cells_num
    fld_host = model.gpu.arrays.cells.get()
    col_host = model.gpu.arrays.colors.get()
    module = SourceModule(model.cuda_source)
    emit, absorb = module.get_function("emit"), module.get_function("absorb")
    fld, col = to_gpu(fld_host), to_gpu(col_host)
    for _ in range(10):
        launch(emit, [fld, ctypes.c_long(n)], n)
        launch(absorb, [fld, col, ctypes.c_long(n)], n)
    out = np.empty_like(fld_host)
    _check(_lib.xen_memcpy_d2h(_vp(out.ctypes.data), fld,
                               ctypes.c_size_t(out.nbytes), None))
    for _ in range(10):
        model.step()
    assert np.array_equal(out[:n], model.gpu.arrays.cells.get()[:n])
    _check(_lib.xen_free(fld))
    _check(_lib.xen_free(col))
    _check(_lib.xen_module_unload(module.mod))
