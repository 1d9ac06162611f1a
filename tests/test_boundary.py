"""The C ABI library loads and exports every symbol its headers declare."""
import ctypes
import glob
import os
import re

from xentica_amd.runtime import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for header in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(header).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for match in re.finditer(r"\b(xen_\w+)\s*\(", text):
            names.add(match.group(1))
    return sorted(names)


def test_headers_declare_functions():
    assert "xen_compile" in declared_functions()
    assert "xen_launch" in declared_functions()


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(hip.library_path())
    missing = [name for name in declared_functions()
               if not hasattr(lib, name)]
    assert not missing, missing


def test_binding_loads_without_gpu():
    lib = hip.library()
    assert lib.xen_last_error() is not None


def test_hiprtc_compiles_generated_source_on_cpu():
    """hiprtc needs no GPU: the generated plain kernels compile here."""
    from examples.game_of_life import GameOfLife, GOLExperiment
    model = GameOfLife(GOLExperiment)
    code = hip.compile_source(model.cuda_source, "test.hip")
    assert code[:4] == b"\x7fELF" or len(code) > 1000


def test_compile_error_is_reported():
    import pytest
    from xentica_amd.core.exceptions import XenticaException
    with pytest.raises(XenticaException, match="hiprtc"):
        hip.compile_source('extern "C" __global__ void k() { nonsense; }',
                           "bad.hip")


def test_device_first_call_does_not_deadlock():
    """``hip.device()`` before anything else loads the library under its
    lock; it must return or raise, never block (no GPU here: it raises)."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from xentica_amd.runtime import hip\n"
            "from xentica_amd.core.exceptions import XenticaException\n"
            "try:\n    hip.device()\nexcept XenticaException:\n    pass\n"
            "print('done')\n") % ROOT
    proc = subprocess.run([sys.executable, "-c", code], capture_output=True,
                          text=True, timeout=60)
    assert "done" in proc.stdout, proc.stderr
