"""
Fused emit+absorb kernels: buffers in LDS, several steps per launch.

For a model on a 2-D orthogonal lattice (Moore, von Neumann or any
orthogonal neighbourhood with unit offsets) with a torus or static
border, the DSL methods are translated a second time with the tile
addressing scheme (``core/addressing.py:TileAddressing``) and wrapped in
one generated kernel:

* a workgroup owns a ``TX x TY`` output tile and loads the region around
  it -- halo ``M = h * T`` where ``T`` is the number of steps per launch
  and ``h`` is 1, or 2 when ``emit`` reads neighbours' main state -- into
  LDS, resolving the torus wrap or the static border while loading;
* each step runs ``emit`` over the cells whose inputs are still valid
  (``distance to the region edge >= s*h + h - 1``), writing the buffer
  slots to LDS planes, then ``absorb``+``color`` over the cells at
  distance ``>= (s+1)*h`` into the other main-state plane, so buffers
  never reach HBM and the main state makes one HBM round trip per ``T``
  steps (the reference writes all ``k`` buffer planes to HBM every step,
  ``xentica/core/base.py:254-283, 461-487``);
* the colours of the output cells live in LDS for the whole launch, so
  the per-step colour effect (``color_effects.py:106-130``) is exact;
  where the colour a step writes provably does not depend on the colour
  before it (``runtime/ranges.py``) and the stored colours' bounds, which
  the host follows, meet the proof's precondition, a second form of the
  kernel computes the colour on a launch's last step only and does not
  read the stored colours;
* the kernel is persistent over a list of (instance, tile) work items
  described by a per-instance table (offsets, sizes, parameter values),
  so one launch advances any number of independent automata of uneven
  sizes and different parameters; a multi-step run is captured as a HIP
  graph.

Step semantics are those of the reference's plain structure: the emit of
a step sees the main state before it, absorb sees every buffer written by
that step's emit, slots emit never writes stay zero (as in the
reference's zero-initialised buffer planes).
"""
import numpy as np

from xentica_amd.core.color_effects import ColorEffect
from xentica_amd.core.exceptions import XenticaException
from xentica_amd.core.topology.border import TorusBorder, StaticBorder
from xentica_amd.core.topology.lattice import OrthogonalLattice
from xentica_amd.core.topology.neighborhood import OrthogonalNeighborhood
from xentica_amd.runtime import hip, ranges

__all__ = ['FusedProgram', 'FusedEngine', 'Batch', 'eligible', 'autotune',
           'render_tables', 'TUNE_TILES', 'TUNE_DEPTHS']


class _DoesNotFit(XenticaException):
    """The requested tile and depth exceed the LDS budget."""

#: LDS one workgroup may use (two workgroups per CU fit in 160 KiB).
LDS_BUDGET = 80 * 1024
BLOCK = 256

_PARAM_FIELD = {'int': ('int', np.int32), 'float': ('float', np.float32),
                'bool': ('int', np.int32)}


def eligible(model):
    """``(True, '')`` if the fused kernels can run ``model``, else a reason."""
    topology = model.topology
    if topology.dimensions != 2:
        return False, "fused kernels are 2-D"
    if not isinstance(topology.lattice, OrthogonalLattice):
        return False, "lattice is not orthogonal"
    hood = topology.neighborhood
    if not isinstance(hood, OrthogonalNeighborhood):
        return False, "neighbourhood is not orthogonal"
    if any(abs(step) > 1 for delta in hood.neighbor_deltas for step in delta):
        return False, "neighbour offsets beyond 1"
    if type(topology.border) not in (TorusBorder, StaticBorder):
        return False, "border is neither TorusBorder nor StaticBorder"
    for param in model._emit_params + model._absorb_params:
        if param.ctype not in _PARAM_FIELD:
            return False, "parameter %s has C type %s" % (param.name,
                                                          param.ctype)
    return True, ""


class FusedProgram:
    """The generated fused kernel of one model class and tile shape."""

    #: (tile, depth) candidates, best first, for states of at most four
    #: bytes and for wider ones; the first that fits the LDS budget wins.
    #: The orders follow what the autotuner measured on MI355X: byte-sized
    #: Life-like states run fastest on 64x16 tiles at 6-8 steps per launch,
    #: 8- and 16-byte words (EvoLife, EnergyLife) on 32x16 tiles at 2.
    CANDIDATES = (((64, 16), 8), ((64, 16), 6), ((32, 32), 4), ((32, 16), 4),
                  ((32, 32), 2), ((32, 16), 2), ((16, 16), 2), ((16, 16), 1),
                  ((8, 8), 1))
    CANDIDATES_WIDE = (((32, 16), 2), ((32, 16), 3), ((16, 16), 2),
                       ((16, 16), 1), ((8, 8), 1))

    def __init__(self, model, tile=None, depth=None, fold=True, split=True):
        if tile is None or depth is None:
            wide = np.dtype(model.main.dtype).itemsize > 4
            for cand_tile, cand_depth in (self.CANDIDATES_WIDE if wide
                                          else self.CANDIDATES):
                try:
                    self.__init__(model, cand_tile, cand_depth, fold, split)
                    return
                except _DoesNotFit:
                    continue
            raise XenticaException("fused engine: no tile fits in LDS")
        ok, why = eligible(model)
        if not ok:
            raise XenticaException("fused engine: " + why)
        emit_body, absorb_body, access = model.translate_tile(fold)
        if access.emit_writes_main:
            raise XenticaException(
                "fused engine: emit() stores the main state (the reference "
                "races on it); use the plain engine")
        if access.emit_reads_neighbor_buffer or access.absorb_writes_buffer:
            raise XenticaException(
                "fused engine: emit reads neighbour buffers or absorb "
                "writes buffers")
        # folded emit: absorb reads buffers as functions of the main words
        self.folded = access.fold is not None
        self.halo_per_step = (2 if access.emit_reads_neighbor_main
                              and not self.folded else 1)
        self.depth = int(depth)
        self.tile = (int(tile[0]), int(tile[1]))
        self.k = len(model.buffers)
        planes = 0 if self.folded else self.k
        self.word = model.main.dtype
        self.ctype = model.main.ctype
        self.params = list({p.name: p for p in model._emit_params
                            + model._absorb_params}.values())
        self.emit_slots = sorted(access.emit_slots)
        self.static = isinstance(model.topology.border, StaticBorder)
        self.border_value = (model.topology.border.value
                             if self.static else 0)
        margin = self.halo_per_step * self.depth
        self.region = (self.tile[0] + 2 * margin, self.tile[1] + 2 * margin)
        area = self.region[0] * self.region[1]
        itemsize = np.dtype(self.word).itemsize
        if area * self.region[0] >= 1 << 22:
            raise _DoesNotFit("fused engine: region %s too large for the "
                              "kernel's row arithmetic" % (self.region, ))
        self.lds_bytes = ((2 + planes) * area * itemsize
                          + 12 * (self.tile[0] * self.tile[1] + BLOCK))
        if self.lds_bytes > LDS_BUDGET:
            raise _DoesNotFit(
                "fused engine: tile %s at depth %d needs %d bytes of LDS" % (
                    self.tile, self.depth, self.lds_bytes))
        self.inst_dtype = self._instance_dtype()
        self.width_prefix = model.topology.lattice.width_prefix
        self._render_fns, self._render_modules = {}, {}
        defines = "".join(
            const.get_define_code() for name, const in model.constants.items()
            if not name.startswith(model.topology.lattice.width_prefix))
        # the main word of a computed cell is copied forward unless absorb
        # always stores it; the colour of halo cells (never written back) is
        # skipped when the colour code leaves the state alone
        self.main_copy = not access.absorb_writes_main
        self.color_guard = access.color_is_pure
        # stored colours from which every step's colour is its new value
        # (None: unknown), and the new values' per-channel bounds
        self.color_plan, self.color_values = _color_plan(model, access)
        # products of an integer 0 or 1 written as selects
        absorb_code = _with_selects(model, access.absorb_code)
        # output tile rows split evenly over the block: column-owning
        # threads with compile-time region offsets (see _STEP_LOOP_SPLIT)
        groups = BLOCK // self.tile[0] if BLOCK % self.tile[0] == 0 else 0
        self.split = (split and self.folded and groups > 0
                      and self.tile[1] % groups == 0)
        self.rows_per_thread = self.tile[1] // groups if self.split else 0
        # cells per ring run unrolled in the kernel (longer runs loop)
        self.ring_run = max(1, min(self.halo_per_step * self.depth - 1, 8))
        if self.color_guard:
            absorb_body = (absorb_code + "if (_inner) {\n"
                           + access.color_code + "}\n" + access.store_code)
        self.source = self._render(defines, emit_body, absorb_body, None,
                                   margin)
        # the pieces of each form, for the render-epilogue variants
        self._forms = {False: (defines, emit_body, absorb_body, None, margin)}
        # the form that computes the colour on a launch's last step only,
        # without reading the stored colour: each step is held twice, with
        # the colour and without it
        self.source_fast = None
        self.lds_bytes_fast = self.lds_bytes
        self.direct_colors = False
        if self.color_plan is not None:
            memory = type(model).color.memory
            color_code = access.color_code.replace(memory, "") if memory \
                else access.color_code
            self._forms[True] = (
                defines, emit_body,
                absorb_code + "if (_inner) {\n" + color_code + "}\n"
                + access.store_code,
                absorb_code + access.store_code, margin)
            # without the colour staging the LDS holds only the planes; the
            # last step then stores each colour itself (a wave's 64 cells
            # of a row are contiguous), which pays where it raises the
            # workgroups a CU holds by half or more
            planes_only = (2 + planes) * area * itemsize
            if (_per_cu(planes_only) >= 1.5 * _per_cu(self.lds_bytes)
                    and color_code.count(_COLOR_STORE) == 1):
                self.direct_colors = True
                self.lds_bytes_fast = planes_only
                color_code = color_code.replace(_COLOR_STORE,
                                                _COLOR_STORE_DIRECT)
            self.source_fast = self._render(
                defines, emit_body,
                absorb_code + "if (_inner) {\n" + color_code + "}\n"
                + access.store_code,
                absorb_code + access.store_code, margin,
                cdirect=self.direct_colors)
        self._module = self._module_fast = None

    def _instance_dtype(self):
        fields = [('cells', np.int64), ('cols', np.int64), ('w0', np.int32),
                  ('w1', np.int32), ('tiles_x', np.int32), ('tile0', np.int32)]
        for param in self.params:
            fields.append(('p_' + param.name, _PARAM_FIELD[param.ctype][1]))
        size = sum(np.dtype(t).itemsize for _, t in fields)
        if size % 8:
            fields.append(('pad', np.int32))
        return np.dtype(fields, align=True)

    def _render(self, defines, emit_body, absorb_body, absorb_plain,
                margin, cdirect=False, render=None):
        """The kernel source; with ``absorb_plain`` (the absorb body
        without the colour) the form that computes the colour only on a
        launch's last step and never reads the stored colours; with
        ``render`` (the per-pixel render body and its extra kernel
        arguments) the form whose tiles also write their viewport pixels
        from the colours in LDS after the last step."""
        tx, ty = self.tile
        render_args = render_epilogue = ""
        if render is not None:
            body, extra = render
            render_args = (", const int *__restrict__ _rt, int *img, "
                           "int zoom, int dx, int dy, int width" + extra)
            render_epilogue = _RENDER_EPILOGUE.format(
                w=self.width_prefix, render_body=body)
        params_struct = "".join(
            "%s p_%s;\n" % (_PARAM_FIELD[p.ctype][0], p.name)
            for p in self.params)
        if self.inst_dtype.names[-1] == 'pad':
            params_struct += "int pad;\n"
        param_locals = "".join(
            "const %s %s = (%s) _in.p_%s;\n" % (p.ctype, p.name, p.ctype, p.name)
            for p in self.params)
        zero_slots = "".join(
            "_tb[%d * _RA + _q] = 0;\n" % (slot - 1)
            for slot in range(1, self.k + 1) if slot not in self.emit_slots)
        # an off-board neighbour reads as the constant in every slot
        static_emit = "".join(
            "_tb[%d * _RA + _li] = (_word) _BVALUE;\n" % (slot - 1)
            for slot in range(1, self.k + 1))
        border = "1" if self.static else "0"
        cfast = 0 if absorb_plain is None else 1
        if self.folded:
            zero_slots = tb_decl = emit_section = ""
        else:
            tb_decl = "__shared__ _word _tb[_K * _RA];\n" if self.k else \
                "_word *_tb = 0;\n"
            emit_section = _EMIT_SECTION.format(
                static_emit=static_emit, emit_body=emit_body)
        main_copy = ("                _mp[_cur ^ 1][_li] = _mp[_cur][_li];"
                     if self.main_copy else "")
        def both(with_color, without_color):
            if absorb_plain is None:
                return with_color
            return ("if (_last) {\n" + with_color + "} else {\n"
                    + without_color + "}\n")

        if self.split:
            def cells(body):
                return "".join(
                    _SPLIT_CELL.format(j=j, main_copy=main_copy,
                                       absorb_body=body)
                    for j in range(self.rows_per_thread))
            run = self.ring_run
            # ring cells never compute the colour
            ring_body = absorb_plain or absorb_body
            step_body = _SPLIT_STEP.format(
                inner_cells=both(cells(absorb_body), cells(absorb_plain)),
                vertical_run=_ring_run("_RW", run, main_copy, ring_body),
                horizontal_run=_ring_run("1", run, main_copy, ring_body))
            step_loop = _STEP_LOOP_SPLIT.format(step_body=step_body)
        else:
            def absorb_pass(body):
                return _GENERIC_PASS.format(main_copy=main_copy,
                                            absorb_body=body)
            step_loop = _STEP_LOOP_GENERIC.format(
                emit_section=emit_section,
                absorb_pass=both(absorb_pass(absorb_body),
                                 absorb_pass(absorb_plain)))
        return _KERNEL_TEMPLATE.format(
            defines=defines, tx=tx, ty=ty, margin=margin,
            k=0 if self.folded else self.k, nt=BLOCK,
            rows=self.rows_per_thread, run=self.ring_run,
            he=self.halo_per_step, ctype=self.ctype, static=border,
            cfast=cfast, cdirect=int(cdirect), helpers=ranges.FDIV_HELPER,
            bvalue="%dULL" % (self.border_value & 0xFFFFFFFFFFFFFFFF),
            params_struct=params_struct, param_locals=param_locals,
            zero_slots=zero_slots, tb_decl=tb_decl, step_loop=step_loop,
            render_args=render_args, render_epilogue=render_epilogue)

    @property
    def function(self):
        """The compiled kernel (hiprtc, cached)."""
        if self._module is None:
            self._module = hip.Module.from_source(self.source, "bsca_fused.hip")
        return self._module.get_function("fused_bsca")

    @property
    def function_fast(self):
        """The kernel that computes the colour on a launch's last step."""
        if self._module_fast is None:
            self._module_fast = hip.Module.from_source(
                self.source_fast, "bsca_fused_fast.hip")
        return self._module_fast.get_function("fused_bsca")

    def render_function(self, model, fast):
        """The form ``fast`` of the kernel with the render epilogue (its
        arguments extended by the render tables, the image, zoom, pan,
        viewport width and the model's render parameters), or None when
        the model's render body is not the plain 2-D projection's."""
        if fast not in self._forms:
            return None
        if fast not in self._render_fns:
            function = None
            body = _epilogue_body(model)
            if body is not None:
                extra = "".join(", %s %s" % (p.ctype, p.name)
                                for p in model._render_params)
                module = hip.Module.from_source(
                    self._render(*self._forms[fast], render=(body, extra)),
                    "bsca_fused_render.hip")
                self._render_modules[fast] = module
                function = module.get_function("fused_bsca")
            self._render_fns[fast] = function
        return self._render_fns[fast]

    def admits(self, bounds):
        """Whether stored colours within ``bounds`` (per-channel ``(min,
        max)``, None if unknown) allow :attr:`function_fast`."""
        plan = self.color_plan
        if plan is None:
            return False
        if plan == ():
            return True
        return bounds is not None and all(
            low <= lo and hi <= high
            for (lo, hi), (low, high) in zip(bounds, plan))

    def bounds_after(self, bounds, fast):
        """Colour bounds after a launch from colours within ``bounds``."""
        if self.color_values is not None and (fast or self.admits(bounds)):
            return self.color_values
        return None

    def kernel_for(self, bounds):
        """``(fast, function)`` to launch on colours within ``bounds``."""
        if self.admits(bounds):
            return True, self.function_fast
        return False, self.function

    def tiles_of(self, size):
        """``(tiles_x, tiles_y)`` covering a field of ``size``."""
        return (-(-size[0] // self.tile[0]), -(-size[1] // self.tile[1]))


_KERNEL_TEMPLATE = r"""
{defines}
#define _TX {tx}
#define _TY {ty}
#define _M {margin}
#define _RW (_TX + 2 * _M)
#define _RH (_TY + 2 * _M)
#define _RA (_RW * _RH)
#define _RP (_NT / _RW)
#define _K {k}
#define _NT {nt}
#define _J {rows}
#define _RUN {run}
#define _HE {he}
#define _STATIC {static}
#define _BVALUE {bvalue}
// 1: only a launch's last step computes the colour, and the stored
// colours are not read (launched only where the colour written provably
// equals the new value at every step); 0: every step computes it
{helpers}#define _CFAST {cfast}
// 1: the last step stores the colours straight to HBM (no LDS staging)
#define _CDIRECT {cdirect}
typedef {ctype} _word;

// The cells of the region at distance >= margin from its edge, one per
// thread in turn: (_lx, _ly) and the region index _li.  The row comes
// from a float reciprocal of the rectangle width, exact while
// _RA * _RW < 2^22 (the host checks): (q + 0.5) / w is at least 0.5 / w
// from an integer and the product's error is below _RA * 2^-23.
#define _RECT_FOR(margin) \
    const int _mg = (margin), _ws = _RW - 2 * _mg, _hs = _RH - 2 * _mg; \
    const float _iw = 1.0f / (float) _ws; \
    for (int _q = threadIdx.x, _ry = 0, _lx = 0, _ly = 0, _li = 0; \
         _q < _ws * _hs && (_ry = (int) (((float) _q + 0.5f) * _iw), \
                            _lx = _mg + (_q - _ry * _ws), _ly = _mg + _ry, \
                            _li = _ly * _RW + _lx, true); \
         _q += blockDim.x)

struct _Inst {{
long long cells;
long long cols;
int w0;
int w1;
int tiles_x;
int tile0;
{params_struct}}};

// element index of in-field cell (x, y) of the current instance: both
// coordinates are in range, so the unsigned 32x32->64-bit product is exact
#define _index(y, x) ((unsigned long long) (unsigned) (y) * (unsigned) _in.w0 \
                      + (unsigned) (x))

// torus wrap; only coordinates outside [0, n) (tiles on a field's edge)
// pay for the division
__device__ __forceinline__ int _wrap(int v, int n) {{
    if ((unsigned) v < (unsigned) n) return v;
    v %= n;
    return v < 0 ? v + n : v;
}}

extern "C" __global__ __launch_bounds__(256)
void fused_bsca(const _word *__restrict__ _src_all, _word *__restrict__ _dst_all,
                int3 *__restrict__ _col_all, const _Inst *__restrict__ _insts,
                const int *__restrict__ _tile_inst, int _n_tiles, int _steps{render_args}) {{
    __shared__ _word _mp[2][_RA];
    {tb_decl}
#if !_CDIRECT
    __shared__ int3 _col[_TX * _TY + 256];
#endif
    // A SIMD issues from its oldest ready wave first: workgroups start at
    // the highest user priority and step down after each quarter of their
    // share of tiles, so the workgroups of a CU finish together instead of
    // the youngest running its last tiles alone.
    __builtin_amdgcn_s_setprio(3);
    const int _share = (_n_tiles - (int) blockIdx.x + (int) gridDim.x - 1) / (int) gridDim.x;
    const int _step_at = _share >= 4 ? _share / 4 : 1;
    int _done = 0, _prio = 3;
    for (int _tile = blockIdx.x; _tile < _n_tiles; _tile += gridDim.x) {{
        const _Inst _in = _insts[_tile_inst[_tile]];
        const int _t = _tile - _in.tile0;
        const int _ox = (_t % _in.tiles_x) * _TX;
        const int _oy = (_t / _in.tiles_x) * _TY;
        const _word *_src = _src_all + _in.cells;
        _word *_dst = _dst_all + _in.cells;
        int3 *_cg = _col_all + _in.cols;
        {param_locals}
#if _RW <= _NT
        {{ // thread (column _c, phase _p) loads region rows _p, _p + _RP, ...,
           // eight loads in flight before the LDS writes; the column's wrap
           // is resolved once per tile
            const int _c = (int) threadIdx.x % _RW, _p = (int) threadIdx.x / _RW;
            if (_p < _RP) {{
#if _STATIC
                const int _gx = _ox - _M + _c;
                const bool _offx = _gx < 0 || _gx >= _in.w0;
#else
                const int _gx = _wrap(_ox - _M + _c, _in.w0);
#endif
                for (int _r0 = _p; _r0 < _RH; _r0 += 8 * _RP) {{
                    _word _v[8];
#pragma unroll
                    for (int _k = 0; _k < 8; ++_k) {{
                        const int _r = _r0 + _k * _RP, _gy = _oy - _M + _r;
                        if (_r < _RH) {{
#if _STATIC
                            const bool _off = _offx || _gy < 0 || _gy >= _in.w1;
                            _v[_k] = _off ? (_word) _BVALUE : _src[_index(_gy, _gx)];
#else
                            _v[_k] = _src[_index(_wrap(_gy, _in.w1), _gx)];
#endif
                        }}
                    }}
#pragma unroll
                    for (int _k = 0; _k < 8; ++_k) {{
                        const int _r = _r0 + _k * _RP;
                        if (_r < _RH) {{
                            const int _q = _r * _RW + _c;
                            _mp[0][_q] = _v[_k];
                            {zero_slots}
                        }}
                    }}
                }}
            }}
        }}
#else
        for (int _q = threadIdx.x; _q < _RA; _q += _NT) {{
            const int _gx = _ox - _M + _q % _RW, _gy = _oy - _M + _q / _RW;
#if _STATIC
            const bool _off = _gx < 0 || _gx >= _in.w0 || _gy < 0 || _gy >= _in.w1;
            _mp[0][_q] = _off ? (_word) _BVALUE : _src[_index(_gy, _gx)];
#else
            _mp[0][_q] = _src[_index(_wrap(_gy, _in.w1), _wrap(_gx, _in.w0))];
#endif
            {zero_slots}
        }}
#endif
#if _CFAST
#elif _NT % _TX == 0
        {{ // colours of the output tile: thread (column _cx, phase _ry) loads
           // rows _ry, _ry + _NT / _TX, ... inside the field, eight at a time
            const int _cx = (int) threadIdx.x % _TX, _gx = _ox + _cx;
            const int _ry = (int) threadIdx.x / _TX;
            if (_gx < _in.w0)
                for (int _r0 = _ry; _r0 < _TY; _r0 += 8 * (_NT / _TX)) {{
                    // three int arrays (an int3 array is not kept in registers)
                    int _c0[8], _c1[8], _c2[8];
#pragma unroll
                    for (int _k = 0; _k < 8; ++_k) {{
                        const int _r = _r0 + _k * (_NT / _TX);
                        if (_r < _TY && _oy + _r < _in.w1) {{
                            const int3 _cv = _cg[_index(_oy + _r, _gx)];
                            _c0[_k] = _cv.x;
                            _c1[_k] = _cv.y;
                            _c2[_k] = _cv.z;
                        }}
                    }}
#pragma unroll
                    for (int _k = 0; _k < 8; ++_k) {{
                        const int _r = _r0 + _k * (_NT / _TX);
                        if (_r < _TY && _oy + _r < _in.w1)
                            _col[_r * _TX + _cx] = make_int3(_c0[_k], _c1[_k], _c2[_k]);
                    }}
                }}
        }}
#else
        for (int _q = threadIdx.x; _q < _TX * _TY; _q += _NT) {{
            const int _gx = _ox + _q % _TX, _gy = _oy + _q / _TX;
            if (_gx < _in.w0 && _gy < _in.w1)
                _col[_q] = _cg[_index(_gy, _gx)];
        }}
#endif
        __syncthreads();
{step_loop}#if _NT % _TX == 0
        {{ const int _cx = (int) threadIdx.x % _TX, _gx = _ox + _cx;
            if (_gx < _in.w0)
                for (int _r = (int) threadIdx.x / _TX; _r < _TY && _oy + _r < _in.w1;
                     _r += _NT / _TX) {{
                    const unsigned long long _g = _index(_oy + _r, _gx);
                    _dst[_g] = _mp[_cur][(_r + _M) * _RW + _cx + _M];
#if !_CDIRECT
                    _cg[_g] = _col[_r * _TX + _cx];
#endif
                }}
        }}
#else
        for (int _q = threadIdx.x; _q < _TX * _TY; _q += _NT) {{
            const int _qx = _q % _TX, _qy = _q / _TX;
            const int _gx = _ox + _qx, _gy = _oy + _qy;
            if (_gx < _in.w0 && _gy < _in.w1) {{
                const unsigned long long _g = _index(_gy, _gx);
                _dst[_g] = _mp[_cur][(_qy + _M) * _RW + _qx + _M];
#if !_CDIRECT
                _cg[_g] = _col[_q];
#endif
            }}
        }}
#endif
{render_epilogue}        __syncthreads();
        if (++_done % _step_at == 0 && _prio > 0) {{
            --_prio;
            if (_prio == 2)
                __builtin_amdgcn_s_setprio(2);
            else if (_prio == 1)
                __builtin_amdgcn_s_setprio(1);
            else
                __builtin_amdgcn_s_setprio(0);
        }}
    }}
}}
"""

# The render epilogue: after the last step the colours of the output tile
# are in LDS (_col), and the tile writes every viewport pixel whose cell it
# holds with the model's render body.  The host's tables list, per cell
# column (row), the pixel columns (rows) that show it -- _rt holds the
# column offsets [w0 + 1], the row offsets [w1 + 1], the pixel columns
# ordered by cell column [width] and the pixel rows ordered by cell row --
# so a tile's pixels are one contiguous run of each list.  The render body
# recomputes the cell of pixel i itself; a cell outside the tile (which the
# tables exclude) would read black, never another tile's LDS.
_RENDER_EPILOGUE = r"""        {{
            const int *_rxo = _rt, *_ryo = _rt + _in.w0 + 1;
            const int *_rxp = _ryo + _in.w1 + 1, *_ryp = _rxp + width;
            const int _px0 = _rxo[_ox];
            const int _npx = _rxo[min(_ox + _TX, _in.w0)] - _px0;
            const int _py0 = _ryo[_oy];
            const int _npy = _ryo[min(_oy + _TY, _in.w1)] - _py0;
            const int {w}0 = _in.w0, {w}1 = _in.w1;
            for (int _q = threadIdx.x; _q < _npx * _npy; _q += _NT) {{
                const int _jy = _q / _npx, _jx = _q - _jy * _npx;
                const unsigned i = (unsigned) _ryp[_py0 + _jy] * (unsigned) width
                                   + (unsigned) _rxp[_px0 + _jx];
{render_body}
            }}
        }}
"""


# The step loop of the general form: one pass per step over the valid
# rectangle, any tile shape, with or without the emit pass.
_STEP_LOOP_GENERIC = r"""        int _cur = 0;
        for (int _s = 0; _s < _steps; ++_s) {{
            const int _da = (_s + 1) * _HE;
            const bool _last = _s + 1 == _steps;
{emit_section}
{absorb_pass}
            __syncthreads();
            _cur ^= 1;
        }}
"""

# The absorb pass of one step of the general form.
_GENERIC_PASS = r"""
            {{ _RECT_FOR(_da) {{
#if _STATIC
                const int _gx = _ox - _M + _lx, _gy = _oy - _M + _ly;
                if (_gx < 0 || _gx >= _in.w0 || _gy < 0 || _gy >= _in.w1) {{
                    _mp[_cur ^ 1][_li] = (_word) _BVALUE;
                    continue;
                }}
#endif
{main_copy}
                const bool _inner = _lx >= _M && _lx < _RW - _M && _ly >= _M && _ly < _RH - _M;
                const int _oi = _inner ? (_ly - _M) * _TX + (_lx - _M)
                                       : _TX * _TY + (int) threadIdx.x;
                {{
#if !_CDIRECT
                    int3 *col = &_col[_oi];
#endif
                    const unsigned i = 0;
{absorb_body}
                }}
            }} }}
"""

# The step loop of a folded program whose tile rows split evenly over the
# block: the steps come in pairs with the ping-pong plane fixed in each
# half (_cur is a constant there), the output tile is computed by threads
# that each own _J cells of one column (region offsets known at compile
# time, neighbour words shared by a column's cells), and only the ring of
# halo cells still valid at this step goes through the generic index
# arithmetic.
_STEP_LOOP_SPLIT = r"""
        int _cur = 0;
        for (int _s = 0; _s < _steps; _s += 2) {{
            {{ const int _cur = 0, _da = (_s + 1) * _HE;
            const bool _last = _s + 1 == _steps;
{step_body}
            }}
            __syncthreads();
            if (_s + 1 >= _steps) break;
            {{ const int _cur = 1, _da = (_s + 2) * _HE;
            const bool _last = _s + 2 == _steps;
{step_body}
            }}
            __syncthreads();
        }}
        _cur = _steps & 1;
"""

# One step of the split loop: the halo ring, then the output tile.
_SPLIT_STEP = r"""
            {{ const int _bw = _M - _da;
            if (_bw > 0) {{
                // the ring of halo cells still valid at this step, as runs
                // of _bw cells: one per column of the top band (rows
                // [_da, _M)) and of the bottom band (rows [_RH - _M,
                // _RH - _da)), then one per output row in the left band
                // (columns [_da, _M)) and in the right band (columns
                // [_RW - _M, _RW - _da)); a run's cells share their
                // neighbour words
                const int _ws = _RW - 2 * _da, _nv = 2 * _ws, _nu = _nv + 2 * _TY;
                for (int _u = threadIdx.x; _u < _nu; _u += _NT) {{
                    if (_u < _nv) {{
                        const bool _low = _u >= _ws;
                        const int _lx = _da + (_low ? _u - _ws : _u);
                        const int _ly0 = _low ? _RH - _M : _da;
                        for (int _r0 = 0; _r0 < _bw; _r0 += _RUN) {{
                            const int _li0 = (_ly0 + _r0) * _RW + _lx, _rn = _bw - _r0;
{vertical_run}
                        }}
                    }} else {{
                        const int _v = _u - _nv;
                        const bool _right = _v >= _TY;
                        const int _ly = _M + (_right ? _v - _TY : _v);
                        const int _lx0 = _right ? _RW - _M : _da;
                        for (int _r0 = 0; _r0 < _bw; _r0 += _RUN) {{
                            const int _li0 = _ly * _RW + _lx0 + _r0, _rn = _bw - _r0;
{horizontal_run}
                        }}
                    }}
                }}
            }} }}
            {{ // output tile: column _cx, rows _cy .. _cy + _J - 1
                const int _cx = (int) threadIdx.x % _TX;
                const int _cy = (int) threadIdx.x / _TX * _J;
                const int _li0 = (_M + _cy) * _RW + _M + _cx;
                const int _oi0 = _cy * _TX + _cx;
#if _STATIC
                // cells past a field's far edge hold the border word; they
                // are computed like the others and then overwritten
                const bool _offx = _ox + _cx >= _in.w0;
#endif
{inner_cells}
            }}
"""

# One cell of the output tile (row _j of the thread's column).
_SPLIT_CELL = r"""
                {{ const int _li = _li0 + {j} * _RW, _oi = _oi0 + {j} * _TX;
                    const bool _inner = true;
{main_copy}
                    {{
    #if !_CDIRECT
                    int3 *col = &_col[_oi];
#endif
                        const unsigned i = 0;
{absorb_body}
                    }}
#if _STATIC
                    if (_offx || _oy + _cy + {j} >= _in.w1)
                        _mp[_cur ^ 1][_li] = (_word) _BVALUE;
#endif
                }}
"""


# One ring cell (``_li`` = ``_li0 + {offset}``); colour goes to the
# thread's spare slot.
_RING_CELL = r"""
                            {{ const int _li = _li0 + {offset};
                                const bool _inner = false;
                                const int _oi = _TX * _TY + (int) threadIdx.x;
{main_copy}
                                {{
#if !_CDIRECT
                                    int3 *col = &_col[_oi];
#endif
                                    const unsigned i = 0;
{absorb_body}
                                }}
#if _STATIC
                                {{ const int _gx = _ox - _M + _li % _RW, _gy = _oy - _M + _li / _RW;
                                if (_gx < 0 || _gx >= _in.w0 || _gy < 0 || _gy >= _in.w1)
                                    _mp[_cur ^ 1][_li] = (_word) _BVALUE; }}
#endif
                            }}
"""


def _ring_run(stride, length, main_copy, absorb_body):
    """Cells ``0 .. min(_rn, length) - 1`` of a ring run; cell ``r`` sits
    inside ``if (_rn > r)`` after cell ``r - 1``, so that the loads of the
    earlier cells dominate it and can be reused."""
    def cell(r):
        return _RING_CELL.format(offset="%d * %s" % (r, stride),
                                 main_copy=main_copy, absorb_body=absorb_body)
    text = ""
    for r in reversed(range(1, length)):
        text = "if (_rn > %d) {\n%s%s}\n" % (r, cell(r), text)
    return cell(0) + text


_EMIT_SECTION = r"""
            {{ const int _de = _s * _HE + (_HE - 1);
            _RECT_FOR(_de) {{
#if _STATIC
                const int _gx = _ox - _M + _lx, _gy = _oy - _M + _ly;
                if (_gx < 0 || _gx >= _in.w0 || _gy < 0 || _gy >= _in.w1) {{
                    {static_emit}
                    continue;
                }}
#endif
                {{
{emit_body}
                }}
            }} }}
            __syncthreads();
"""


def _color_plan(model, access):
    """
    When a launch may compute the colour on its last step only.

    Returns ``(plan, values)``.  ``plan`` is None when every step must
    (the colour code changes the state, or nothing is known of its
    values); ``()`` when the colour written never reads the stored one; or
    one ``(low, high)`` pair per channel: from stored colours within these
    the colour of every step is its new value, so only the last step's
    counts.  ``values`` bounds the new values per channel (None if not
    known); they lie within the plan's intervals again.

    For ``MovingAverage`` the pairs are ``[hi - FADE_IN, lo + FADE_OUT]``
    where ``[lo, hi]`` bounds the channel's new value: from a stored colour
    in that interval the clamp returns the new value.
    """
    nothing = (None, None)
    if not access.color_is_pure:
        return nothing
    effect = getattr(type(model), "color", None)
    if not isinstance(effect, ColorEffect):
        return nothing
    if not effect.memory:
        return (), None
    code = access.color_code
    if code.count(effect.memory) != 1:
        return nothing
    fades = []
    for name in ("FADE_IN", "FADE_OUT"):
        const = model.constants.get(name)
        value = None if const is None else const.value
        if not isinstance(value, int) or isinstance(value, bool):
            return nothing
        fades.append(value)
    fade_in, fade_out = fades
    bounds = ranges.channel_bounds(
        access.absorb_code + code[:code.index(effect.memory)],
        *_range_context(model), channels=("new_r", "new_g", "new_b"))
    if bounds is None:
        return nothing
    plan = []
    for lo, hi in bounds:
        if hi - lo > fade_in or hi - lo > fade_out:
            return nothing
        low, high = hi - fade_in, lo + fade_out
        # the clamp's old - FADE_OUT and old + FADE_IN stay in int range
        if low - fade_out < -(1 << 31) or high + fade_in >= 1 << 31:
            return nothing
        plan.append((low, high))
    return tuple(plan), tuple(bounds)


def _with_selects(model, code):
    """
    ``code`` with every product of integers one of which is provably 0 or
    1 written as a select: one ``v_cndmask`` where a 32-bit multiply (the
    DSL's idiom ``flag * value``) is a quarter-rate instruction.  The
    values are those of the C expressions (``runtime/ranges.py``).
    """
    text, _ = ranges.Analysis(*_range_context(model)).rewrite(code)
    return text


def _range_context(model):
    """Names, arrays and typedefs the tile code of ``model`` refers to."""
    analysis = ranges.Analysis()
    names = {}
    for name, const in model.constants.items():
        if isinstance(const.value, bool) or not isinstance(
                const.value, (int, float)):
            continue
        try:
            names[name] = analysis.evaluate(str(const.value))
        except (ranges.Unknown, ranges.GiveUp):
            continue
    params = {p.name: p for p in model._emit_params + model._absorb_params}
    for name, param in params.items():
        ctype = ranges.CTYPES.get((param.ctype, ))
        if ctype is not None:
            names[name] = ranges.Value.full(ctype)
    int_type = ranges.CTYPES[("int", )]
    for name in ("_li", "_cur", "_RW", "_RH", "_RA", "_TX", "_TY", "_M"):
        names[name] = ranges.Value.full(int_type)
    names["i"] = ranges.Value(ranges.CTYPES[("unsigned", )], 0, 0)
    word = ranges.type_named(model.main.ctype)
    types = {"_word": word} if word is not None else {}
    pointers = {"_mp": (word, 2), "_tb": (word, 1)} if word is not None \
        else {}
    return names, pointers, types


def _union(bounds_list):
    """Per-channel bounds covering every entry (None if one is unknown)."""
    if not bounds_list or any(b is None for b in bounds_list):
        return None
    return tuple((min(b[c][0] for b in bounds_list),
                  max(b[c][1] for b in bounds_list))
                 for c in range(len(bounds_list[0])))


def _per_cu(lds_bytes):
    """Workgroups of 256 threads a CU holds with ``lds_bytes`` of LDS each
    (at most 8: 32 waves)."""
    return max(1, min(8, (160 * 1024) // max(1, lds_bytes)))


def _persistent_grid(n_tiles, function):
    """Workgroups of a persistent launch: as many as the CUs hold at once
    (a workgroup that waits for a free CU would finish a whole share of
    tiles after the others)."""
    per_cu, cus = function.occupancy(BLOCK)
    return max(1, min(n_tiles, max(1, per_cu) * max(1, cus)))


_COLOUR_LOAD = "int3 c = col[ii];\n"
# the epilogue's colour of cell (x0, x1): the tile's LDS colours
_COLOUR_LDS = (
    "int3 c = ((unsigned) (x0 - _ox) < (unsigned) _TX\n"
    "          && (unsigned) (x1 - _oy) < (unsigned) _TY)\n"
    "    ? _col[(x1 - _oy) * _TX + (x0 - _ox)] : make_int3(0, 0, 0);\n")


def _epilogue_body(model):
    """The model's per-pixel render body reading the tile's LDS colours,
    or None unless the renderer is the plain projection onto axes (0, 1)
    of a 2-D field (``renderers.py:220-270``)."""
    from xentica_amd.core.renderers import RendererPlain
    renderer = model.renderer
    if type(renderer) is not RendererPlain \
            or tuple(renderer.projection_axes) != (0, 1) \
            or model.topology.dimensions != 2:
        return None
    body = model.kernel_bodies.get('render', "")
    if body.count(_COLOUR_LOAD) != 1:
        return None
    return body.replace(_COLOUR_LOAD, _COLOUR_LDS)


def render_tables(size, width, height, zoom, dx, dy):
    """
    The render epilogue's pixel lists of a viewport: ``int32`` column
    offsets ``[w0 + 1]``, row offsets ``[w1 + 1]``, pixel columns ordered
    by the cell column they show ``[width]`` and pixel rows ordered by cell
    row ``[height]``, concatenated.  A pixel shows the cell of the plain
    renderer's body: ``(int) ((float) p / (float) zoom) + shift``, wrapped
    onto the torus as ``renderers.py:220-270`` does.
    """
    parts_off, parts_pix = [], []
    for count, shift, cells in ((width, dx, size[0]), (height, dy, size[1])):
        pix = np.arange(count, dtype=np.int64)
        cell = (pix.astype(np.float32) / np.float32(zoom)).astype(np.int64)
        cell += shift
        neg = cell < 0
        cell[neg] = cells - (-cell[neg]) % cells
        big = cell >= cells
        cell[big] = cell[big] % cells
        order = np.argsort(cell, kind="stable")
        parts_off.append(np.searchsorted(cell[order], np.arange(cells + 1)))
        parts_pix.append(pix[order])
    return np.concatenate(parts_off + parts_pix).astype(np.int32)


# the colour effect's store, and the last-step form's store straight to
# HBM (output-tile cell _oi of the workgroup's tile, inside the field only)
_COLOR_STORE = "col[i] = make_int3(new_r, new_g, new_b);"
_COLOR_STORE_DIRECT = (
    "{ const int _gxo = _ox + _oi % _TX, _gyo = _oy + _oi / _TX;\n"
    "if (_gxo < _in.w0 && _gyo < _in.w1)\n"
    "    _cg[_index(_gyo, _gxo)] = make_int3(new_r, new_g, new_b); }")


class _Instances:
    """Device tables of a set of model instances run by one program."""

    def __init__(self, program, models, per_instance=True):
        self.program = program
        self.models = list(models)
        self.per_instance = per_instance
        itemsize = np.dtype(program.word).itemsize
        offsets, col_offsets, tile_inst = [], [], []
        cells = cols = 0
        tile0 = 0
        table = np.zeros(len(self.models), dtype=program.inst_dtype)
        for idx, model in enumerate(self.models):
            if len(model.size) != 2:
                raise XenticaException("fused engine: 2-D models only")
            tiles_x, tiles_y = program.tiles_of(model.size)
            table[idx]['cells'] = cells
            table[idx]['cols'] = cols
            table[idx]['w0'], table[idx]['w1'] = model.size
            table[idx]['tiles_x'] = tiles_x
            table[idx]['tile0'] = tile0
            offsets.append(cells)
            col_offsets.append(cols)
            tile_inst += [idx] * (tiles_x * tiles_y)
            tile0 += tiles_x * tiles_y
            cells += model.cells_num
            cols += model.cells_num
        self.table = table
        self.offsets = offsets
        self.col_offsets = col_offsets
        self.total_cells = cells
        self.n_tiles = tile0
        self.itemsize = itemsize
        self.table_buf = hip.Buffer(max(1, table.nbytes))
        self.tiles_buf = hip.Buffer(4 * max(1, self.n_tiles))
        self.tiles_buf.upload(np.array(tile_inst, dtype=np.int32))
        self._uploaded = None

    def param_values(self):
        """Per-instance parameter values (instance overrides first)."""
        values = []
        for model in self.models:
            own = getattr(model, "_param_values", {}) if self.per_instance \
                else {}
            values.append(tuple(own.get(p.name, p.value)
                                for p in self.program.params))
        return values

    def refresh(self, values=None, stream=None):
        """Upload the table when parameter values changed.

        The copy is queued on ``stream``, the stream the table's readers
        run on, and waits for it: a launch still reading the old table
        there finishes first (the null stream does not wait for the
        non-blocking streams batches run on).
        """
        if values is None:
            values = self.param_values()
        if values != self._uploaded:
            for idx, vals in enumerate(values):
                for param, val in zip(self.program.params, vals):
                    self.table[idx]['p_' + param.name] = val
            self.table_buf.upload(self.table, stream)
            self._uploaded = values


class FusedEngine:
    """Single-model engine on the fused kernel, with deferred steps."""

    name = "fused"

    def __init__(self, model, tile=None, depth=None, fold=True, split=True):
        self.model = model
        self.program = FusedProgram(model, tile=tile, depth=depth, fold=fold,
                                    split=split)
        self._inst = None
        self._pending = 0
        self._pending_values = None
        self._other = None
        self._in_other = False
        self._launches = {}
        self._tables = (None, None)
        #: frames rendered by a launch's epilogue (no separate render pass)
        self.epilogue_frames = 0
        arrays = model.gpu.arrays
        arrays.cells.before_read = self.flush
        arrays.colors.before_read = self.flush

    def _setup(self):
        if self._inst is None:
            # a single model reads the class-wide parameter values, as the
            # reference's step() does (xentica/core/base.py:717, 722)
            self._inst = _Instances(self.program, [self.model],
                                    per_instance=False)
            nbytes = self.model.cells_num * self._inst.itemsize
            self._other = hip.Buffer(nbytes)

    def launch_elementwise(self, kernel, args, count):
        """Grid-stride launch (render, through the plain kernels)."""
        from xentica_amd.runtime.engines import elementwise_grid, BLOCK as EB
        kernel(*args, grid=(elementwise_grid(count), 1, 1), block=(EB, 1, 1))

    def step(self):
        """Queue one step with the current parameter values."""
        if self._inst is None:
            self._setup()
        # the class-wide values (what _Instances.param_values() gives for a
        # single model), read without its per-instance lookups
        params = self.program.params
        values = [tuple([p.value for p in params])] if params else [()]
        if self._pending and values != self._pending_values:
            self.flush()
        self._pending_values = values
        self._pending += 1
        if self._pending >= self.program.depth:
            self.flush()

    def _run(self, steps, values, render=None):
        """One launch of ``steps`` steps; ``render``: ``(fast, function,
        extra arguments)`` of the render-epilogue form for this launch."""
        self._inst.refresh(values)
        arrays = self.model.gpu.arrays
        plane0 = arrays.cells.plane_ptr(self.model.cells_num)
        colors = arrays.colors.device_ptr()
        src, dst = ((self._other.ptr, plane0) if self._in_other
                    else (plane0, self._other.ptr))
        fast, function = self.program.kernel_for(arrays.colors.bounds)
        # launches prepared once per (direction, step count, arrays, form)
        key = (src, dst, colors, steps, fast)
        launch = self._launches.get(key)
        if render is not None:
            fast, function, extra = render
            grid = _persistent_grid(self._inst.n_tiles, function)
            launch = function.prepare(
                np.uint64(src), np.uint64(dst), np.uint64(colors),
                np.uint64(self._inst.table_buf.ptr),
                np.uint64(self._inst.tiles_buf.ptr),
                np.int32(self._inst.n_tiles), np.int32(steps), *extra,
                grid=(grid, 1, 1), block=(BLOCK, 1, 1))
        elif launch is None:
            grid = _persistent_grid(self._inst.n_tiles, function)
            launch = function.prepare(
                np.uint64(src), np.uint64(dst), np.uint64(colors),
                np.uint64(self._inst.table_buf.ptr),
                np.uint64(self._inst.tiles_buf.ptr),
                np.int32(self._inst.n_tiles), np.int32(steps),
                grid=(grid, 1, 1), block=(BLOCK, 1, 1))
            if len(self._launches) > 64:
                self._launches.clear()
            self._launches[key] = launch
        launch()
        arrays.colors.bounds = self.program.bounds_after(
            arrays.colors.bounds, fast)
        self._in_other = not self._in_other

    def flush(self):
        """Run the queued steps; leave the state in the cells array."""
        while self._pending:
            steps = min(self._pending, self.program.depth)
            self._run(steps, self._pending_values)
            self._pending -= steps
        if self._in_other:
            nbytes = self.model.cells_num * self._inst.itemsize
            hip.check(hip.library().xen_memcpy_d2d(
                self.model.gpu.arrays.cells.plane_ptr(self.model.cells_num),
                self._other.ptr,
                nbytes, None))
            self._in_other = False

    def render_frame(self, render_params, pixels):
        """
        With steps queued, run them with the render epilogue: the last
        launch's tiles also write the viewport's pixels from the colours
        they hold in LDS, with the model's render body
        (``renderers.py:220-270``), and no separate pass reads the colours
        back.  Returns False (the caller renders from the arrays) when no
        step is queued, the zoom is below 1 or the renderer is not the plain
        2-D projection.
        """
        model = self.model
        if not self._pending or model.zoom < 1:
            return False
        self._setup()
        arrays = model.gpu.arrays
        while self._pending > self.program.depth:
            self._run(self.program.depth, self._pending_values)
            self._pending -= self.program.depth
        fast, _ = self.program.kernel_for(arrays.colors.bounds)
        function = self.program.render_function(model, fast)
        if function is None:
            return False
        args = model.renderer.get_args_vals(model)
        zoom, dx, dy, width = (int(v) for v in args[2:6])
        view = (zoom, dx, dy, width, pixels // width)
        if self._tables[0] != view:
            table = render_tables(model.size, width, pixels // width,
                                  zoom, dx, dy)
            buf = hip.Buffer(table.nbytes)
            buf.upload(table)
            self._tables = (view, buf)
        extra = [np.uint64(self._tables[1].ptr),
                 np.uint64(arrays.img.device_ptr()), np.int32(zoom),
                 np.int32(dx), np.int32(dy), np.int32(width)]
        extra += list(render_params)
        self._run(self._pending, self._pending_values,
                  render=(fast, function, extra))
        self._pending = 0
        self.epilogue_frames += 1
        return True

    def load_state(self, cells, colors):
        """Replace the state (queued steps are applied first)."""
        self.flush()
        arrays = self.model.gpu.arrays
        arrays.cells.set(cells)
        arrays.colors.set(colors)


class Batch:
    """
    Many independent models of one class stepped together.

    Every model keeps its own size, seed and parameter values (set through
    ``model.<param> = value`` or its experiment); one persistent launch
    advances all of them ``depth`` steps.  ``run(steps)`` captures the
    launch sequence in a HIP graph and replays it.

    :param models: instances of one ``CellularAutomaton`` subclass
    :param tile: output tile of a workgroup
    :param depth: steps per launch
    """

    def __init__(self, models, tile=(32, 32), depth=4, stream=None,
                 fold=True, split=True):
        models = list(models)
        if not models:
            raise XenticaException("empty batch")
        cls = type(models[0])
        if any(type(m) is not cls for m in models):
            raise XenticaException("a batch holds instances of one class")
        self.program = FusedProgram(models[0], tile=tile, depth=depth,
                                    fold=fold, split=split)
        self.models = models
        self.stream = stream or hip.Stream()
        self.inst = _Instances(self.program, models)
        total = self.inst.total_cells * self.inst.itemsize
        self.state = [hip.Buffer(total), hip.Buffer(total)]
        self.colors = hip.Buffer(12 * self.inst.total_cells)
        self.color_bounds = None
        self.cur = 0
        self._graphs = {}
        self.gather()

    def gather(self):
        """Copy every model's current state into the batch buffers."""
        from xentica_amd.runtime.hip import check, library
        lib = library()
        for model in self.models:
            model._engine.flush()
        # the flushes queue kernels and copies on the null stream, which
        # the batch's non-blocking stream does not wait for; the batch's
        # own earlier launches read the buffers the copies overwrite
        hip.device().synchronize()
        for model, off, coff in zip(self.models, self.inst.offsets,
                                    self.inst.col_offsets):
            arrays = model.gpu.arrays
            check(lib.xen_memcpy_d2d(
                self.state[self.cur].ptr + off * self.inst.itemsize,
                arrays.cells.plane_ptr(model.cells_num),
                model.cells_num * self.inst.itemsize, self.stream.handle))
            check(lib.xen_memcpy_d2d(
                self.colors.ptr + coff * 12, arrays.colors.device_ptr(),
                model.cells_num * 12, self.stream.handle))
        self.stream.synchronize()
        self.color_bounds = _union(
            [model.gpu.arrays.colors.bounds for model in self.models])

    def scatter(self):
        """Copy the batch state back into every model's arrays."""
        from xentica_amd.runtime.hip import check, library
        lib = library()
        for model, off, coff in zip(self.models, self.inst.offsets,
                                    self.inst.col_offsets):
            arrays = model.gpu.arrays
            check(lib.xen_memcpy_d2d(
                arrays.cells.plane_ptr(model.cells_num),
                self.state[self.cur].ptr + off * self.inst.itemsize,
                model.cells_num * self.inst.itemsize, self.stream.handle))
            check(lib.xen_memcpy_d2d(
                arrays.colors.device_ptr(), self.colors.ptr + coff * 12,
                model.cells_num * 12, self.stream.handle))
            arrays.colors.bounds = self.color_bounds
            if hasattr(model._engine, "invalidate"):
                model._engine.invalidate()
        self.stream.synchronize()

    def save(self, filename):
        """Write every instance's state in the reference's checkpoint
        format (``xentica/core/base.py:755-763``): a pickled
        ``{"instances": [{"cells", "colors", "random"}, ...]}`` in the
        order of :attr:`models`; each entry is what that model's own
        ``save()`` writes."""
        from xentica_amd.runtime import checkpoint
        self.synchronize()
        self.scatter()
        checkpoint.write(filename, {"instances": [
            {"cells": model.gpu.arrays.cells.get(),
             "colors": model.gpu.arrays.colors.get(),
             "random": model.random} for model in self.models]})

    def load(self, filename):
        """Read a checkpoint written by :meth:`save` into the models and
        the batch (``xentica/core/base.py:765-771`` per instance)."""
        from xentica_amd.runtime import checkpoint
        states = checkpoint.read(filename)["instances"]
        if len(states) != len(self.models):
            raise XenticaException("checkpoint holds %d instances, the batch "
                                   "%d" % (len(states), len(self.models)))
        self.synchronize()
        for model, state in zip(self.models, states):
            model._engine.load_state(state["cells"], state["colors"])
            model.random.load(state["random"])
        self.gather()

    def _launch(self, steps, fast):
        function = (self.program.function_fast if fast
                    else self.program.function)
        grid = _persistent_grid(self.inst.n_tiles, function)
        function(
            np.uint64(self.state[self.cur].ptr),
            np.uint64(self.state[1 - self.cur].ptr),
            np.uint64(self.colors.ptr), np.uint64(self.inst.table_buf.ptr),
            np.uint64(self.inst.tiles_buf.ptr), np.int32(self.inst.n_tiles),
            np.int32(steps), grid=(grid, 1, 1), block=(BLOCK, 1, 1),
            stream=self.stream)
        self.cur = 1 - self.cur

    def launch_plan(self, steps):
        """Steps per launch of a run of ``steps`` steps."""
        depth = self.program.depth
        return [depth] * (steps // depth) + ([steps % depth] if steps % depth
                                              else [])

    def run(self, steps, graph=True):
        """Advance every instance ``steps`` steps (asynchronous; a change
        of parameter values waits for the batch's earlier launches)."""
        self.inst.refresh(stream=self.stream)
        fast, function = self.program.kernel_for(self.color_bounds)
        plan = self.launch_plan(steps)
        if not plan:
            return
        bounds = self.program.bounds_after(self.color_bounds, fast)
        if not graph:
            for count in plan:
                self._launch(count, fast)
            self.color_bounds = bounds
            return
        key = (tuple(plan), self.cur, fast)
        if key not in self._graphs:
            start = self.cur
            hip.Graph.begin(self.stream)
            try:
                for count in plan:
                    self._launch(count, fast)
            finally:
                graph_obj = hip.Graph.end(self.stream)
            self._graphs[key] = (graph_obj, self.cur)
            self.cur = start
        graph_obj, end = self._graphs[key]
        graph_obj.launch(self.stream)
        self.cur = end
        self.color_bounds = bounds

    def synchronize(self):
        """Wait for the batch's stream."""
        self.stream.synchronize()


#: Tile shapes and temporal depths the autotuner tries (those that fit the
#: LDS budget for the model's word size and fold).
TUNE_TILES = ((32, 32), (64, 16), (32, 16), (64, 32), (128, 8), (16, 16))
TUNE_DEPTHS = (1, 2, 3, 4, 6, 8)


def autotune(models, trial_steps=64, tiles=TUNE_TILES, depths=TUNE_DEPTHS,
             stream=None, log=None):
    """
    Pick the tile shape and temporal depth of a batch by measurement.

    Every ``(tile, depth)`` pair whose kernel fits the LDS budget is
    compiled, run once to warm up and then timed with HIP events over
    ``trial_steps`` steps (rounded up to whole launches, graph-captured
    as in :meth:`Batch.run`) on the models' current states; the models
    themselves are not advanced.  Returns ``(batch, timings)``: the
    fastest :class:`Batch`, ready to run from the models' states, and
    ``{(tile, depth): cell-updates per second}``.
    """
    models = list(models)
    stream = stream or hip.Stream()
    cells = sum(model.cells_num for model in models)
    timings = {}
    best = None
    for tile in tiles:
        for depth in depths:
            try:
                FusedProgram(models[0], tile=tile, depth=depth)
            except _DoesNotFit:
                continue
            batch = Batch(models, tile=tile, depth=depth, stream=stream)
            steps = -(-trial_steps // depth) * depth
            batch.run(steps)
            start, stop = hip.Event(), hip.Event()
            start.record(stream)
            batch.run(steps)
            stop.record(stream)
            stop.synchronize()
            rate = cells * steps / (start.elapsed_ms(stop) / 1e3)
            timings[(tile, depth)] = rate
            if log:
                log("autotune tile %s depth %d: %.4g cell-updates/s"
                    % (tile, depth, rate))
            if best is None or rate > best[0]:
                best = (rate, tile, depth)
            del batch
    if best is None:
        raise XenticaException("fused engine: no tile fits in LDS")
    _, tile, depth = best
    return Batch(models, tile=tile, depth=depth, stream=stream), timings
