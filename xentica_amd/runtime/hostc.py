"""
Host C++ rendering of a model's generated kernel bodies, built with g++
and OpenMP.

The DSL translator produces C statements (``core/base.py``); the GPU
engines wrap them in gfx950 kernels.  This module wraps the same
statements in host functions, for two uses:

* :class:`PlainHost` -- the reference's plain structure on the CPU (an
  ``emit`` loop over every cell writing the buffer planes, then an
  ``absorb`` + colour loop reading them, ``xentica/core/base.py:254-283,
  461-499, 703-728``), with ``#pragma omp parallel for`` over cells.  It
  is the CPU baseline of the benchmark: the reference semantics executed
  on the host's cores, for any model the DSL expresses.
* :class:`LifeLikeProbe` -- the tile form of ``absorb`` + ``color`` (the
  statements the fused kernels run, with neighbours as offsets in a
  region) evaluated on *every* configuration of a cell's 3^d block.  For
  a model whose state is one bit on a Moore torus this decides exactly
  whether its step is a Life-like rule ``next = born(count) |
  sustain(count) & state`` with a colour that depends on the new state
  only, and which rule -- the condition for running it on the bit-sliced
  kernels.

Nothing here is a fallback for the GPU engines: the probe only decides
which GPU engine runs a model, and :class:`PlainHost` only measures.
Compiled libraries are cached on disk by a hash of their source.
"""
import ctypes
import hashlib
import os
import subprocess
import threading

import numpy as np

from xentica_amd.core.exceptions import XenticaException

__all__ = ['PRELUDE', 'compile_host', 'PlainHost', 'LifeLikeProbe',
           'host_threads']

#: Host definitions of what the generated statements use from HIP.
PRELUDE = r"""
#include <math.h>
#include <stdint.h>
#include <string.h>
struct int3 { int x, y, z; };
static inline int3 make_int3(int x, int y, int z) { int3 v; v.x = x; v.y = y; v.z = z; return v; }
static inline int min(int a, int b) { return a < b ? a : b; }
static inline int max(int a, int b) { return a > b ? a : b; }
static inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
static inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
static inline long long min(long long a, long long b) { return a < b ? a : b; }
static inline long long max(long long a, long long b) { return a > b ? a : b; }
static inline float min(float a, float b) { return fminf(a, b); }
static inline float max(float a, float b) { return fmaxf(a, b); }
static inline int __popc(unsigned v) { return __builtin_popcount(v); }
static inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
"""

#: g++ flags: host CPU, OpenMP, C semantics for floats (no contraction).
CXXFLAGS = ("-O3", "-march=native", "-fopenmp", "-ffp-contract=off",
            "-shared", "-fPIC", "-w", "-std=c++17")

_PARAM_CTYPES = {'int': ctypes.c_int, 'float': ctypes.c_float,
                 'bool': ctypes.c_bool, 'unsigned int': ctypes.c_uint}

_LOCK = threading.Lock()
_LOADED = {}


def host_threads():
    """Cores this process may run on."""
    return len(os.sched_getaffinity(0))


def _cache_dir():
    root = os.environ.get("XENTICA_CACHE_DIR") or os.path.join(
        os.path.expanduser("~"), ".cache", "xentica_amd")
    path = os.path.join(root, "hostc")
    os.makedirs(path, exist_ok=True)
    return path


def compile_host(source, compiler="g++"):
    """Compile ``source`` (C++) into a shared library; returns its CDLL."""
    key = hashlib.sha256(
        ("\0".join(CXXFLAGS) + "\0" + source).encode()).hexdigest()[:32]
    with _LOCK:
        if key in _LOADED:
            return _LOADED[key]
        path = os.path.join(_cache_dir(), key + ".so")
        if not os.path.exists(path):
            src = os.path.join(_cache_dir(), "%s.%d.cpp" % (key, os.getpid()))
            with open(src, "w") as out:
                out.write(source)
            tmp = "%s.%d.tmp" % (path, os.getpid())
            try:
                proc = subprocess.run(
                    [compiler, *CXXFLAGS, src, "-o", tmp],
                    capture_output=True, text=True)
            except OSError as err:
                raise XenticaException("host compiler unavailable: %s" % err)
            finally:
                if os.path.exists(src):
                    os.remove(src)
            if proc.returncode:
                raise XenticaException(
                    "host compile failed:\n" + proc.stderr[-4000:])
            os.replace(tmp, path)
        lib = ctypes.CDLL(path)
        _LOADED[key] = lib
        return lib


def _param_ctype(param):
    try:
        return _PARAM_CTYPES[param.ctype]
    except KeyError:
        raise XenticaException("host code: parameter %s has C type %s"
                               % (param.name, param.ctype))


def _host_loop(name, args, body):
    """A per-cell host function: the grid-stride kernel's loop as an
    OpenMP loop over ``i`` (``core/base.py:_elementwise_kernel``)."""
    arg_string = ", ".join("%s %s" % (ctype, var) for ctype, var in args)
    return (
        'extern "C" void %s(%s, long n) {\n'
        "#pragma omp parallel for schedule(static)\n"
        "for (long _ii = 0; _ii < n; ++_ii) {\n"
        "const unsigned i = (unsigned) _ii;\n"
        "%s"
        "}\n"
        "}\n\n") % (name, arg_string, body)


class PlainHost:
    """
    The model's plain emit/absorb bodies as host functions over the
    reference's layout: ``fld`` = ``n * (k + 1) + 1`` container words
    (plane 0 the main state, plane ``p`` buffer slot ``p``) and ``col`` =
    ``int32[3n]``.  ``step()`` runs ``emit`` then ``absorb`` with the
    parameter values the model's class holds, as ``step()`` of the
    reference does (``xentica/core/base.py:703-728``).
    """

    def __init__(self, model):
        self.model = model
        bodies = model.kernel_bodies
        ctype = model.main.ctype
        self.emit_params = list(model.emit_params)
        self.absorb_params = list(model.absorb_params)
        emit_args = [(ctype, "*fld")] + [(p.ctype, p.name)
                                         for p in self.emit_params]
        absorb_args = [(ctype, "*fld"), ("int3", "*col")] + [
            (p.ctype, p.name) for p in self.absorb_params]
        self.source = (PRELUDE + model.build_defines()
                       + _host_loop("emit", emit_args, bodies['emit'])
                       + _host_loop("absorb", absorb_args,
                                    bodies['absorb']))
        lib = compile_host(self.source)
        self._emit = lib.emit
        self._emit.argtypes = ([ctypes.c_void_p]
                               + [_param_ctype(p) for p in self.emit_params]
                               + [ctypes.c_long])
        self._emit.restype = None
        self._absorb = lib.absorb
        self._absorb.argtypes = ([ctypes.c_void_p, ctypes.c_void_p]
                                 + [_param_ctype(p)
                                    for p in self.absorb_params]
                                 + [ctypes.c_long])
        self._absorb.restype = None
        self._omp = _openmp_control(lib)
        n = model.cells_num
        self.fld = np.zeros(n * (len(model.buffers) + 1) + 1,
                            dtype=model.main.dtype)
        self.col = np.zeros(3 * n, dtype=np.int32)

    def load(self, cells, colors):
        """Copy a state (reference layout) and colours in."""
        self.fld[:] = np.asarray(cells, dtype=self.fld.dtype).reshape(-1)
        self.col[:] = np.asarray(colors, dtype=np.int32).reshape(-1)

    def step(self, steps=1, threads=0):
        """Advance ``steps`` steps on ``threads`` cores (0: all usable)."""
        self._omp(threads or host_threads())
        n = self.model.cells_num
        for _ in range(steps):
            self._emit(self.fld.ctypes.data,
                       *[p.value for p in self.emit_params], n)
            self._absorb(self.fld.ctypes.data, self.col.ctypes.data,
                         *[p.value for p in self.absorb_params], n)

    @property
    def cells(self):
        """The main-state plane."""
        return self.fld[:self.model.cells_num]


def _openmp_control(lib):
    """``omp_set_num_threads`` of the library's OpenMP runtime."""
    func = lib.omp_set_num_threads
    func.argtypes = [ctypes.c_int]
    func.restype = None
    return func


_PROBE_TEMPLATE = r"""
{prelude}
{defines}
#define _RW 3
#define _RH 3
#define _RA {area}
typedef {ctype} _word;

static inline void _eval(unsigned long long _cfg, {param_args}const int3 _old,
                         unsigned char *_next, int3 *_colour) {{
    _word _mp[2][_RA];
    for (int _p = 0; _p < _RA; ++_p) {{
        _mp[0][_p] = (_word) ((_cfg >> _p) & 1ULL);
        _mp[1][_p] = 0;
    }}
    const int _cur = 0, _li = {centre};
    const bool _inner = true;
    const unsigned i = 0;
    int3 _c = _old;
    int3 *col = &_c;
    _mp[1][_li] = _mp[0][_li];
    {{
{body}
    }}
    *_next = (unsigned char) _mp[1][_li];
    *_colour = _c;
}}

extern "C" void probe_eval(const unsigned long long *cfgs, long m, {param_args}
                           int o0, int o1, int o2, unsigned char *next, int *colour) {{
    const int3 old = make_int3(o0, o1, o2);
    for (long q = 0; q < m; ++q) {{
        int3 c;
        _eval(cfgs[q], {param_names}old, &next[q], &c);
        colour[3 * q] = c.x;
        colour[3 * q + 1] = c.y;
        colour[3 * q + 2] = c.z;
    }}
}}

// Every configuration of the block: the number whose new state differs
// from table[state][count] or whose colour differs from colours[new state]
// (count over the neighbour positions, state at the centre).
extern "C" long long probe_verify({param_args}int o0, int o1, int o2,
                                  const unsigned char *table, const int *colours,
                                  const int *positions, int k,
                                  unsigned long long *first_bad) {{
    const int3 old = make_int3(o0, o1, o2);
    const unsigned long long total = 1ULL << _RA;
    long long bad = 0;
    unsigned long long first = ~0ULL;
#pragma omp parallel for schedule(static) reduction(+ : bad) reduction(min : first)
    for (long long q = 0; q < (long long) total; ++q) {{
        const unsigned long long cfg = (unsigned long long) q;
        int count = 0;
        for (int j = 0; j < k; ++j) count += (int) ((cfg >> positions[j]) & 1ULL);
        const int state = (int) ((cfg >> {centre}) & 1ULL);
        unsigned char next;
        int3 c;
        _eval(cfg, {param_names}old, &next, &c);
        const int want = table[state * (k + 1) + count];
        const bool ok = want == next && next <= 1 &&
                        c.x == colours[3 * next] && c.y == colours[3 * next + 1] &&
                        c.z == colours[3 * next + 2];
        if (!ok) {{
            ++bad;
            if (cfg < first) first = cfg;
        }}
    }}
    *first_bad = first;
    return bad;
}}
"""


class LifeLikeProbe:
    """
    Exact recognition of a Life-like model from its generated code.

    Construction checks the structure: a 2-D or 3-D orthogonal lattice,
    the Moore neighbourhood, a torus, a state of one 1-bit property that
    ``emit`` copies into every buffer, and a colour the fused engine may
    compute on a launch's last step only (``runtime/fused.py``).  It
    raises :class:`XenticaException` otherwise.

    :meth:`rule` then evaluates the model's own absorb+colour statements
    on all ``2^(3^d)`` configurations of a cell's block for given
    interactive parameter values, and returns ``(born, sustain, colours)``
    when every one of them gives ``next = born(count) | sustain(count) &
    state`` (``count`` over the neighbourhood's offsets: eight in 2-D; in
    3-D the reference's list of 26, centre kept and ``(1, 1, 1)``
    dropped, ``xentica/core/topology/neighborhood.py:159-167``) and the
    colour ``colours[next]``; else ``None``.
    """

    def __init__(self, model):
        from xentica_amd.core.topology.border import TorusBorder
        from xentica_amd.core.topology.lattice import OrthogonalLattice
        from xentica_amd.core.topology.neighborhood import MooreNeighborhood
        from xentica_amd.runtime.fused import _color_plan
        topology = model.topology
        dims = topology.dimensions
        if dims not in (2, 3):
            raise XenticaException("bit-sliced engine: 2-D or 3-D only")
        if type(topology.lattice) is not OrthogonalLattice \
                or type(topology.neighborhood) is not MooreNeighborhood \
                or type(topology.border) is not TorusBorder:
            raise XenticaException(
                "bit-sliced engine: orthogonal lattice, Moore neighbourhood "
                "and torus only")
        props = list(model.main.properties.values())
        if len(props) != 1 or props[0].bit_width != 1 or \
                np.dtype(model.main.dtype).itemsize != 1:
            raise XenticaException("bit-sliced engine: one 1-bit property")
        for param in model.emit_params + model.absorb_params:
            _param_ctype(param)
        _, body, access = model.translate_tile(fold=True)
        slots = len(model.buffers)
        if access.fold is None or any(
                access.fold.get(slot) != (1, 0)
                for slot in range(1, slots + 1)):
            raise XenticaException(
                "bit-sliced engine: emit must copy the state into every "
                "buffer")
        if not access.color_is_pure:
            raise XenticaException("bit-sliced engine: the colour changes "
                                   "the state")
        plan, values = _color_plan(model, access)
        if plan is None:
            raise XenticaException(
                "bit-sliced engine: the colour depends on earlier colours")
        self.plan, self.color_values = plan, values
        self.dims = dims
        self.params = list({p.name: p for p in model.emit_params
                            + model.absorb_params}.values())
        deltas = topology.neighborhood.neighbor_deltas
        strides = [1, 3, 9][:dims]
        self.centre = sum(strides)
        self.positions = [self.centre + sum(d * s for d, s in zip(delta,
                                                                   strides))
                          for delta in deltas]
        self.k = len(self.positions)
        defines = "".join(
            const.get_define_code() for name, const in model.constants.items()
            if not name.startswith(model.topology.lattice.width_prefix))
        param_args = "".join("%s %s, " % (p.ctype, p.name)
                             for p in self.params)
        param_names = "".join("%s, " % p.name for p in self.params)
        self.source = _PROBE_TEMPLATE.format(
            prelude=PRELUDE, defines=defines, area=3 ** dims,
            ctype=model.main.ctype, centre=self.centre,
            param_args=param_args, param_names=param_names, body=body)
        lib = compile_host(self.source)
        ptypes = [_param_ctype(p) for p in self.params]
        self._eval = lib.probe_eval
        self._eval.argtypes = ([ctypes.c_void_p, ctypes.c_long] + ptypes
                               + [ctypes.c_int] * 3
                               + [ctypes.c_void_p, ctypes.c_void_p])
        self._eval.restype = None
        self._verify = lib.probe_verify
        self._verify.argtypes = (ptypes + [ctypes.c_int] * 3
                                 + [ctypes.c_void_p] * 3
                                 + [ctypes.c_int, ctypes.c_void_p])
        self._verify.restype = ctypes.c_longlong
        self._omp = _openmp_control(lib)
        self._rules = {}

    def values(self):
        """The parameter values a single model's step reads: the class's,
        as the reference's ``step()`` passes them (``base.py:717, 722``)."""
        if not self.params:
            return ()
        return tuple([p.value for p in self.params])

    def _stored_colour(self):
        """A stored colour from which every step's colour is its new value."""
        if not self.plan:
            return (0, 0, 0)
        return tuple(int(low) for low, _ in self.plan)

    def _canonical(self):
        """One block configuration per reachable (state, count)."""
        configs = {}
        centre_listed = self.centre in self.positions
        others = [p for p in self.positions if p != self.centre]
        for state in (0, 1):
            base = (1 << self.centre) if state else 0
            first = state if centre_listed else 0
            for extra in range(len(others) + 1):
                cfg = base
                for pos in others[:extra]:
                    cfg |= 1 << pos
                configs[(state, first + extra)] = cfg
        return configs

    def rule(self, values):
        """``(born, sustain, colours)`` for these parameter values, or
        ``None`` if the model is not that Life-like rule for them."""
        try:
            return self._rules[values]
        except (KeyError, TypeError):  # new values, or not a tuple
            values = tuple(values)
        if values in self._rules:
            return self._rules[values]
        result = self._probe(values)
        self._rules[values] = result
        return result

    def _probe(self, values):
        args = [_param_ctype(p)(float(v) if p.ctype == 'float' else int(v))
                for p, v in zip(self.params, values)]
        old = self._stored_colour()
        canon = self._canonical()
        keys = sorted(canon)
        cfgs = np.array([canon[key] for key in keys], dtype=np.uint64)
        nxt = np.zeros(len(keys), dtype=np.uint8)
        col = np.zeros(3 * len(keys), dtype=np.int32)
        self._eval(cfgs.ctypes.data, len(keys), *args, *old,
                   nxt.ctypes.data, col.ctypes.data)
        if np.any(nxt > 1):
            return None
        table = np.full(2 * (self.k + 1), 2, dtype=np.uint8)
        colours = {}
        for idx, (state, count) in enumerate(keys):
            table[state * (self.k + 1) + count] = nxt[idx]
            colour = tuple(int(v) for v in col[3 * idx:3 * idx + 3])
            if colours.setdefault(int(nxt[idx]), colour) != colour:
                return None
        born = sustain = 0
        for count in range(self.k + 1):
            dead = table[count]
            alive = table[self.k + 1 + count]
            if dead == 1:
                born |= 1 << count
                if alive == 0:
                    # a birth count that kills live cells: not of the form
                    # born | sustain & state
                    return None
            if alive == 1:
                sustain |= 1 << count
        coltab = np.zeros(6, dtype=np.int32)
        for state in (0, 1):
            coltab[3 * state:3 * state + 3] = colours.get(
                state, colours.get(1 - state))
        positions = np.array(self.positions, dtype=np.int32)
        first = ctypes.c_ulonglong(0)
        self._omp(host_threads())
        bad = self._verify(*args, *old, table.ctypes.data, coltab.ctypes.data,
                           positions.ctypes.data, self.k, ctypes.byref(first))
        if bad:
            return None
        return born, sustain, (tuple(int(v) for v in coltab[:3]),
                               tuple(int(v) for v in coltab[3:]))
