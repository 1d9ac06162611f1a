"""
The base class of all models and the DSL -> HIP translator.

Models subclass :class:`CellularAutomaton`, declare properties and an
inner ``Topology`` class, and describe the buffered-state (BSCA) step in
``emit()``, ``absorb()`` and ``color()``.  When a model is instantiated
those methods are *executed* against recording proxies and every
operation appends a C statement to the kernel body, exactly as the
reference does (``xentica/core/base.py:239-524``).  The bodies are then
wrapped in kernels for gfx950 and compiled with hiprtc through the
project's own C ABI (``include/xentica_hip.h``).

Step semantics (``xentica/core/base.py:703-728``): ``emit`` runs over all
cells, then ``absorb`` (with ``color``) runs over all cells; interactive
parameters are read at every step.  How a step is executed is chosen by
the model's engine (``xentica_amd.runtime.engines``): the ``plain`` engine
keeps the reference's structure with every buffer plane in HBM and is the
parity bridge for any model the DSL can express.
"""
import functools
import itertools
import operator
import pickle
import threading
from copy import deepcopy

import numpy as np

from xentica_amd.bridge import MoireBridge
from xentica_amd.core import _context
from xentica_amd.core.addressing import PlainAddressing, TileAddressing
from xentica_amd.core.properties import Property, ContainerProperty, UINT128
from xentica_amd.core.renderers import RendererPlain
from xentica_amd.core.exceptions import XenticaException
from xentica_amd.runtime import arrays as gpuarray

__all__ = ['BSCA', 'CellularAutomaton', 'CachedNeighbor', 'Translator',
           'GPUKernels', 'GPUArrays', 'GPU']


class CachedNeighbor:
    """The main and buffered state of one neighbour."""

    def __init__(self):
        """Initialize the empty main and buffered states."""
        self.main = ContainerProperty()
        self.buffer = ContainerProperty()


class BSCA(type):
    """
    Metaclass of :class:`CellularAutomaton`.

    Builds the ``main``, ``buffers`` and ``neighbors`` containers of a
    model class and the pairing of buffer slots (``base.py:116-236``).
    """

    mandatory_fields = ('dimensions', 'lattice', 'neighborhood', 'border')

    def __new__(mcs, name, bases, attrs):
        """Build a new model class."""
        attrs['__ordered__'] = [
            key for key in attrs.keys()
            if key not in ('__module__', '__qualname__')]
        new_class = super().__new__(mcs, name, bases, attrs)
        if not any(isinstance(base, BSCA) for base in bases):
            return new_class
        mcs._prepare_topology(new_class, attrs)
        new_class.main = None
        new_class.buffers = None
        new_class.neighbors = None
        mcs._prepare_properties(new_class, bases, attrs)
        new_class.constants = {}
        if not hasattr(new_class, 'renderer'):
            new_class.renderer = RendererPlain()
        new_class.size = tuple(1 for _ in range(new_class.topology.dimensions))
        return new_class

    def _prepare_topology(cls, attrs):
        """Copy and wire the topology objects (``base.py:165-186``)."""
        if hasattr(cls, 'Topology'):
            attrs['Topology'] = cls.Topology
        cls.topology = attrs.get('Topology')
        if cls.topology is None:
            raise XenticaException("No Topology class declared.")
        for field in BSCA.mandatory_fields:
            if not hasattr(cls.topology, field):
                raise XenticaException(
                    "No %s declared in Topology class." % field)
        topology = cls.topology
        topology.lattice = deepcopy(topology.lattice)
        topology.neighborhood = deepcopy(topology.neighborhood)
        topology.border = deepcopy(topology.border)
        topology.lattice.dimensions = topology.dimensions
        topology.neighborhood.dimensions = topology.dimensions
        topology.border.dimensions = topology.dimensions
        topology.neighborhood.topology = topology
        topology.border.topology = topology

    def _prepare_properties(cls, bases, attrs):
        """
        Build the containers and the buffer-slot pairing.

        Properties come from the direct bases' ``__dict__`` and then the
        class itself; a redefinition keeps its slot (``base.py:188-221``).
        ``buffers[q]`` is stored in slot ``rev[q] + 1``;
        ``neighbors[j].buffer`` reads slot ``j + 1`` of neighbour ``j``;
        ``neighbors[j].main`` reads plane 0 of it (``base.py:223-236``).
        """
        cls.main = ContainerProperty()
        cls.buffers = []
        cls.neighbors = []
        cls.meta = type('MetaParams', (object,), {})
        num_neighbors = len(cls.topology.neighborhood)
        for _ in range(num_neighbors):
            cls.buffers.append(ContainerProperty())
            cls.neighbors.append(CachedNeighbor())
        items = [base.__dict__.items() for base in bases]
        items.append(attrs.items())
        restricted = {"main", "buffer"}
        for obj_name, obj in itertools.chain.from_iterable(items):
            if obj_name in restricted:
                continue
            if isinstance(obj, Property):
                cls.main[obj_name] = deepcopy(obj)
                cls.main[obj_name].var_name = "_cell_%s" % obj_name
                for i in range(num_neighbors):
                    cls.buffers[i][obj_name] = deepcopy(obj)
                    cls.buffers[i][obj_name].var_name = "_bcell_%s%d" % (
                        obj_name, i)
                    neighbor = cls.neighbors[i]
                    neighbor.main[obj_name] = deepcopy(obj)
                    neighbor.main[obj_name].var_name = "_dcell_%s%d" % (
                        obj_name, i)
                    neighbor.buffer[obj_name] = deepcopy(obj)
                    neighbor.buffer[obj_name].var_name = "_dbcell_%s%d" % (
                        obj_name, i)
            elif obj.__class__.__name__ == "Parameter":
                obj.name = obj_name
                setattr(cls.meta, obj_name, obj)

        cls.main.buf_num = 0
        cls.main.nbr_num = -1
        cls.main.var_name = "_cell"
        rev = cls.topology.neighborhood.rev_index_map
        for i in range(num_neighbors):
            cls.buffers[i].buf_num = rev[i] + 1
            cls.buffers[i].nbr_num = -1
            cls.buffers[i].var_name = "_bcell%i" % i
            cls.neighbors[i].main.buf_num = 0
            cls.neighbors[i].main.nbr_num = i
            cls.neighbors[i].main.var_name = "_dcell%d" % i
            cls.neighbors[i].buffer.buf_num = i + 1
            cls.neighbors[i].buffer.nbr_num = i
            cls.neighbors[i].buffer.var_name = "_dbcell%d" % i


def _elementwise_kernel(name, args, body):
    """
    Wrap a per-cell body in a grid-stride kernel (``base.py:254-283``).

    The cell count ``n`` is a ``long`` so that plane offsets ``n * p`` do
    not overflow on large fields; cell indices stay ``unsigned``.
    """
    arg_string = ", ".join("%s %s" % (ctype, var) for ctype, var in args)
    return (
        'extern "C" __global__ void %s(%s, long n) {\n'
        "unsigned tid = threadIdx.x;\n"
        "unsigned total_threads = gridDim.x * blockDim.x;\n"
        "unsigned cta_start = blockDim.x * blockIdx.x;\n"
        "for (unsigned i = cta_start + tid; i < n; i += total_threads) {\n"
        "%s"
        "}\n"
        "}\n\n"
    ) % (name, arg_string, body)


class _IdentitySet:
    """A set of DSL objects compared by identity."""

    def __init__(self):
        self._items = {}

    def add(self, obj):
        self._items[id(obj)] = obj

    def __contains__(self, obj):
        return id(obj) in self._items

    def __len__(self):
        return len(self._items)


class Translator:
    """Python -> C translation by execution (``base.py:239-524``)."""

    def __init__(self):
        self.cuda_source = ""
        self._func_body = []
        self._deferred_writes = {}
        self._declarations = _IdentitySet()
        self._unpacks = _IdentitySet()
        self._params = {}
        self._emit_params = []
        self._absorb_params = []
        self._render_params = []
        self._coords_declared = False
        self.addressing = PlainAddressing(self)
        self.kernel_bodies = {}

    def _translate_code(self, *funcs, marks=None):
        """
        Run the DSL methods and return the generated body.

        ``marks``, if a list, receives one ``(code pieces, deferred writes,
        declarations, unpacks, property assignments)`` count tuple before
        each method and one at the end of the methods (before the deferred
        stores).
        """
        self._func_body = []
        self._deferred_writes = {}
        self._declarations = _IdentitySet()
        self._unpacks = _IdentitySet()
        self._params = {}
        self._coords_declared = False

        def mark():
            if marks is not None:
                marks.append((len(self._func_body), len(self._deferred_writes),
                              len(self._declarations), len(self._unpacks),
                              getattr(self.addressing, "assign_count", 0)))

        with _context.translating(self):
            for func in funcs:
                mark()
                func()
            mark()
            # stores happen at the end of the body, in first-write order
            for prop in list(self._deferred_writes.values()):
                prop.deferred_write()
        return "".join(self._func_body)

    def append_code(self, code):
        """Append ``code`` to the kernel body."""
        self._func_body.append(code)

    def deferred_write(self, prop):
        """Schedule the packed store of a container at the end of the body."""
        self._deferred_writes.setdefault(id(prop), prop)

    def declare(self, prop):
        """Mark a variable declared."""
        self._declarations.add(prop)

    def unpack(self, prop):
        """Mark a container unpacked."""
        self._unpacks.add(prop)

    def is_declared(self, prop):
        """Test if a variable is declared."""
        return prop in self._declarations

    def is_unpacked(self, prop):
        """Test if a container is unpacked."""
        return prop in self._unpacks

    def declare_coords(self):
        """Mark coordinate variables declared."""
        self._coords_declared = True

    @property
    def coords_declared(self):
        """Test if coordinate variables are declared."""
        return self._coords_declared

    def define_constant(self, constant):
        """Register a ``#define``."""
        self.constants[constant.name] = deepcopy(constant)

    def is_constant(self, constant_name):
        """Test if a ``#define`` is registered."""
        return constant_name in self.constants

    def define_parameter(self, param):
        """Register an interactive parameter as a kernel argument."""
        self._params[param.name] = param

    def is_parameter(self, param_name):
        """Test if an interactive parameter is registered."""
        return param_name in self._params

    def build_source(self):
        """Generate the emit, absorb and render kernels and the defines."""
        source = self.build_emit()
        source += self.build_absorb()
        source += self.build_render()
        self.cuda_source = self.build_defines() + source

    @property
    def hip_source(self):
        """The generated HIP source of the plain kernels."""
        return self.cuda_source

    def build_defines(self):
        """``#define`` lines of every registered constant."""
        return "".join(const.get_define_code()
                       for const in self.constants.values())

    def build_emit(self):
        """Generate the ``emit`` kernel."""
        body = self._translate_code(self.emit)
        self.kernel_bodies['emit'] = body
        self._emit_params = list(self._params.values())
        args = [(self.main.ctype, "*fld")]
        args += [(param.ctype, param.name) for param in self._emit_params]
        return _elementwise_kernel("emit", args, body)

    def build_absorb(self):
        """Generate the ``absorb`` kernel (absorb and colour)."""
        body = self._translate_code(self.absorb, self.color)
        self.kernel_bodies['absorb'] = body
        self._absorb_params = list(self._params.values())
        args = [(self.main.ctype, "*fld"), ("int3", "*col")]
        args += [(param.ctype, param.name) for param in self._absorb_params]
        return _elementwise_kernel("absorb", args, body)

    def build_render(self):
        """Generate the ``render`` kernel."""
        with _context.translating(self):
            body = self.renderer.render_code()
        self.kernel_bodies['render'] = body
        return _elementwise_kernel("render", self.renderer.args, body)

    def translate_tile(self, fold=True):
        """
        Translate emit and absorb+color for the fused LDS-tile kernels.

        Returns ``(emit_body, absorb_body, addressing)``; the addressing
        object records the access pattern the kernel must honour.  The
        plain source is left untouched.  ``fold=False`` keeps buffer reads
        on the buffer planes even when emit could be folded.
        """
        plain = self.addressing
        tile = TileAddressing(self)
        self.addressing = tile
        marks = []
        try:
            tile.phase = "emit"
            emit_body = self._translate_code(self.emit)
            tile.start_absorb(fold)
            self._translate_code(self.absorb, self.color, marks=marks)
            pieces = list(self._func_body)
            tile.phase = None
        finally:
            self.addressing = plain
        (c_code, c_defer, c_decl, c_unp, c_asg), \
            (e_code, e_defer, e_decl, e_unp, e_asg) = marks[1:]
        tile.absorb_code = "".join(pieces[:c_code])
        tile.color_code = "".join(pieces[c_code:e_code])
        tile.store_code = "".join(pieces[e_code:])
        # the colour may be skipped where it is discarded only if it leaves
        # the state alone: no assignment to a field (a generator advance is
        # one), no new stores, no first unpack or declaration of a field
        # that the stores after it would then use
        tile.color_is_pure = (e_asg == c_asg and e_defer == c_defer
                              and e_unp == c_unp
                              and not any(
                                  isinstance(obj, Property)
                                  for obj in list(
                                      self._declarations._items.values()
                                  )[c_decl:e_decl]))
        return emit_body, "".join(pieces), tile

    def index_to_coord(self, i):
        """Coordinates of flat index ``i`` (or an index array)."""
        return self.topology.lattice.index_to_coord(i, self)

    def pack_state(self, state):
        """Pack ``{name: value}`` into container words, without masking."""
        if self.main.dtype is UINT128:
            return _pack_wide(self.main.properties, state)
        val = 0
        shift = 0
        for name, prop in self.main.properties.items():
            if name in state:
                val += state[name] << shift
            shift += prop.bit_width
        return val


def _pack_wide(properties, state):
    """
    :meth:`pack_state` for 128-bit words: ``sum(value << shift)`` modulo
    2**128 with each value sign-extended from int64, as the 64-bit path
    computes modulo 2**64; returns a :data:`UINT128` array.
    """
    size = max(np.size(val) for val in state.values()) if state else 1
    lo = np.zeros(size, dtype=np.uint64)
    hi = np.zeros(size, dtype=np.uint64)
    shift = 0
    with np.errstate(over='ignore'):
        for name, prop in properties.items():
            if name in state:
                val = np.broadcast_to(
                    np.asarray(state[name], dtype=np.int64), (size, ))
                v_lo = val.astype(np.uint64)
                v_hi = np.where(val < 0, np.uint64(2 ** 64 - 1),
                                np.uint64(0))
                if 0 < shift < 64:
                    v_hi = (v_hi << np.uint64(shift)) | (
                        v_lo >> np.uint64(64 - shift))
                    v_lo = v_lo << np.uint64(shift)
                elif shift >= 64:
                    v_hi = v_lo << np.uint64(shift - 64)
                    v_lo = np.zeros_like(v_lo)
                new_lo = lo + v_lo
                hi = hi + v_hi + (new_lo < lo).astype(np.uint64)
                lo = new_lo
            shift += prop.bit_width
    out = np.empty(size, dtype=UINT128)
    out['lo'], out['hi'] = lo, hi
    return out


class GPUKernels:
    """The compiled plain kernels of a model (``base.py:527-541``)."""

    def __init__(self, source, fp_contract="off"):
        """Compile ``source`` for gfx950 and look the kernels up."""
        from xentica_amd.runtime import hip
        module = hip.Module(hip.compile_source(
            source, "bsca_plain.hip", hip.compile_options(fp_contract)))
        self.module = module
        self.emit = module.get_function("emit")
        self.absorb = module.get_function("absorb")
        self.render = module.get_function("render")


class GPUArrays:
    """
    The device arrays of a model (``base.py:544-569``): ``cells`` in the
    reference's layout (``n * (k + 1) + 1`` words, of which the seeds write
    the main plane and the rest starts at zero) and ``colors`` (``3 n``
    ints, zero at the start).
    """

    def __init__(self, init_plane, cells_total, num_cells):
        """Hold the cells (main plane given, buffer planes zero) and the
        colours (zero)."""
        self.img = None
        self.colors = gpuarray.zeros((3 * num_cells, ), dtype=np.int32)
        self.colors.track_bounds(3)
        self.cells = gpuarray.to_gpu(init_plane, cells_total)

    def init_img(self, num_cells):
        """Allocate the render image."""
        self.img = gpuarray.zeros((num_cells, ), dtype=np.int32)


class GPU:
    """Kernels and arrays of a model; kernels compile on first use."""

    def __init__(self, source, init_plane, cells_total, num_cells,
                 fp_contract="off"):
        """Keep the source; allocate nothing on the device yet."""
        self._source = source
        self._fp_contract = fp_contract
        self._kernels = None
        self.arrays = GPUArrays(init_plane, cells_total, num_cells)

    @property
    def kernels(self):
        """The compiled plain kernels."""
        if self._kernels is None:
            self._kernels = GPUKernels(self._source, self._fp_contract)
        return self._kernels


class CellularAutomaton(Translator, metaclass=BSCA):
    """
    The base class for all models.

    Builds the kernels' source at construction, seeds the field and keeps
    the device state.  Then ``step()`` advances one step, ``render()``
    projects the colours to an RGB image::

        ca = MyCA(MyExperiment)
        ca.set_viewport((320, 200))
        for _ in range(100):
            ca.step()
        frame = ca.render()

    :param experiment_class: the ``Experiment`` holding sizes, seeds and
        parameter values.
    """

    #: Step engine: ``"auto"`` (the bit-sliced engine for Life-like
    #: models on a 1-bit state, else the fused LDS-tile engine when the
    #: model qualifies, else ``"plain"``), ``"bits"``, ``"fused"``,
    #: ``"plain"`` (the reference's structure), ``"fused_or_plain"``, or
    #: ``"fused_generic"``/``"fused_unfolded"`` (other forms of the fused
    #: kernel, for checking); an experiment may override it.
    engine = "auto"
    #: Floating-point contraction of the generated code: ``"off"`` keeps
    #: every ``a * b + c`` as two roundings, ``"fast"`` lets the compiler
    #: fuse them (the default of the reference's nvcc); an experiment may
    #: override it.
    fp_contract = "off"

    def __init__(self, experiment_class):
        """Translate the model, seed the field and prepare the GPU state."""
        super().__init__()
        with _context.translating(self):
            self._init_from_experiment(experiment_class)

    def _init_from_experiment(self, experiment_class):
        self.frame_buf = np.zeros((3, ), dtype=np.uint8)
        self.width, self.height = 0, 0
        # imported here: the seeds import the DSL's expressions, whose
        # package imports this module
        from xentica_amd.seeds.random import LocalRandom
        self.random = LocalRandom(experiment_class.word)
        experiment_class.seed.random = self.random
        experiment_class.random = self.random
        for attr_name in dir(experiment_class):
            attr = getattr(experiment_class, attr_name)
            if callable(attr) or attr_name.startswith("__"):
                continue
            if attr_name == 'seed':
                continue
            if hasattr(self.meta, attr_name):
                self.meta.__dict__[attr_name].__set__(self, attr)
                continue
            setattr(self, attr_name, attr)
        self.speed = 1
        self.paused = False
        self.timestep = 0
        self.cells_num = functools.reduce(operator.mul, self.size)
        self.build_source()
        cells_total = self.cells_num * (len(self.buffers) + 1) + 1
        # seeds write the main plane; the buffer planes start at zero and
        # are allocated only where an engine uses them
        init_plane = np.zeros((self.cells_num, ), dtype=self.main.dtype)
        experiment_class.seed.generate(init_plane, self)
        self.gpu = GPU(self.cuda_source, init_plane, cells_total,
                       self.cells_num, self.fp_contract)
        self.bridge = MoireBridge
        self.renderer.setup_actions(self.bridge)
        self._lock = threading.Lock()
        self._img_bytes = None  # device uint8 copy of the render image
        from xentica_amd.runtime import engines
        self._engine = engines.create(self, self.engine)

    @property
    def emit_params(self):
        """Interactive parameters passed to ``emit``, in argument order."""
        return list(self._emit_params)

    @property
    def absorb_params(self):
        """Interactive parameters passed to ``absorb``, in argument order."""
        return list(self._absorb_params)

    def flush(self):
        """Finish steps the engine has deferred (reads do this implicitly)."""
        self._engine.flush()

    def apply_speed(self, dval):
        """Change the simulation speed (never below 1)."""
        self.speed = max(1, (self.speed + dval))

    def toggle_pause(self):
        """Toggle the ``paused`` flag; a paused model does not step."""
        self.paused = not self.paused

    def set_viewport(self, size):
        """Set the viewport size and allocate the image."""
        self.width, self.height = size
        self.gpu.arrays.init_img(self.width * self.height * 3)

    def step(self):
        """Advance one step; ``timestep`` counts the steps taken."""
        if self.paused:
            return
        with self._lock:
            self._engine.step()
            self.timestep += 1

    def render(self):
        """Render the current colours as ``uint8`` RGB, ``w * h * 3`` values."""
        if self.gpu.arrays.img is None:
            msg = "Viewport is not set, call set_viewport() before rendering."
            raise XenticaException(msg)
        with self._lock:
            pixels = self.width * self.height
            render_params = [param.dtype(param.value)
                             for param in self._render_params]
            # an engine that keeps the state in its own form may project it
            # into the image itself (the bits engine reads the packed cells)
            direct = getattr(self._engine, "render_frame", None)
            if direct is None or not direct(render_params, pixels):
                self._engine.flush()
                args = self.renderer.get_args_vals(self) + render_params
                args.append(np.int64(pixels))
                self._engine.launch_elementwise(
                    self.gpu.kernels.render, args, pixels)
            # the reference's img.get().astype(uint8), with the cast on the
            # device: a quarter of the bytes cross to the host
            from xentica_amd.runtime import hip, kernels
            img = self.gpu.arrays.img
            count = int(np.prod(img.shape))
            if self._img_bytes is None or self._img_bytes.nbytes < count:
                self._img_bytes = hip.Buffer(count)
            kernels.low_bytes(img.device_ptr(), self._img_bytes.ptr, count)
            frame = np.empty(count, dtype=np.uint8)
            self._img_bytes.download(frame)
            return frame

    def save(self, filename):
        """Save cells, colours and the random streams (``base.py:755-763``)."""
        with open(filename, "wb") as ca_file:
            ca_state = {
                "cells": self.gpu.arrays.cells.get(),
                "colors": self.gpu.arrays.colors.get(),
                "random": self.random,
            }
            pickle.dump(ca_state, ca_file)

    def load(self, filename):
        """Load a state written by :meth:`save` (``base.py:765-771``)."""
        with open(filename, "rb") as ca_file:
            ca_state = pickle.load(ca_file)
        with self._lock:
            self._engine.load_state(ca_state['cells'], ca_state['colors'])
            self.random.load(ca_state['random'])
