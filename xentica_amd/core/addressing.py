"""
Addressing schemes: where a container's packed word lives in memory.

Properties ask ``bsca.addressing`` for the C lvalue of a container and for
the declaration code of a neighbour's container; everything else about
code generation is shared by all kernels.

:class:`PlainAddressing` renders the reference's buffered-state layout
verbatim: one flat array ``fld`` of ``n * (k + 1) + 1`` words, plane 0 the
main state and plane ``p`` buffer slot ``p`` (``xentica/core/base.py:659-661``,
``properties.py:217-233``); neighbours are reached through coordinate
arithmetic with wrap or static-border handling (``properties.py:431-494``).
"""

import re

__all__ = ['PlainAddressing', 'TileAddressing', ]

_INT_LITERAL = re.compile(r"^\(?\s*-?\d+\s*\)?$")


class PlainAddressing:
    """The reference's ``fld[i + n * plane]`` layout."""

    name = "plain"

    def __init__(self, model):
        self._model = model

    def cell_ref(self, container):
        """C lvalue of ``container``'s word for the current cell ``i``."""
        if container.nbr_num >= 0:
            neighborhood = self._model.topology.neighborhood
            return neighborhood.neighbor_state(
                container.nbr_num, container.buf_num, "_nx")
        if container.buf_num > 0:
            return "fld[i + n * %d]" % container.buf_num
        return "fld[i]"

    def store_ref(self, container):
        """C lvalue the container's packed word is stored to."""
        return self.cell_ref(container)

    def note_assignment(self, prop, value):
        """Assignments are not tracked by the plain layout."""

    def neighbor_declaration(self, container):
        """Coordinates, border handling and the load of a neighbour word."""
        model = self._model
        topology = model.topology
        lattice, border = topology.lattice, topology.border
        code = ""
        if not model.coords_declared:
            code += lattice.index_to_coord_code("i", "_x")
            code += "int %s;\n" % ", ".join(
                "_nx%d" % axis for axis in range(topology.dimensions))
            model.declare_coords()
        code += topology.neighborhood.neighbor_coords(
            container.nbr_num, "_x", "_nx")
        off_board = lattice.is_off_board_code("_nx")
        if hasattr(border, "wrap_coords"):
            code += "if (%s) {\n%s}\n" % (off_board, border.wrap_coords("_nx"))
            if container.init_val is None:
                code += "%s %s;\n" % (container.ctype, container.var_name)
            else:
                code += "%s %s = %s;\n" % (
                    container.ctype, container.var_name,
                    self.cell_ref(container))
            return code
        var = container.var_name
        code += "%s %s;\nif (%s) {\n%s = %s;\n} else {\n%s = %s;\n}\n" % (
            container.ctype, var, off_board,
            var, border.off_board_state("_nx"),
            var, self.cell_ref(container))
        return code


class TileAddressing:
    """
    LDS tile layout of the fused kernels.

    A workgroup holds a region of the field (output tile plus halo) in
    LDS: two main-state planes ``_mp[2][_RA]`` (read plane ``_cur``,
    written plane ``_cur ^ 1``) and ``_K`` buffer planes ``_tb[_K * _RA]``;
    ``_li`` is the cell's index in the region and ``_RW``/``_RH`` its
    width and height.  Neighbours are fixed offsets in the region: the
    kernel resolved torus wrapping or the static border when it loaded the
    region, so no coordinate code is generated.

    While translating, it records what the model does in each phase
    (``phase`` is set by the translator to ``"emit"`` or ``"absorb"``):
    which buffer slots emit stores, whether emit stores the main state or
    reads neighbours' main state, and whether absorb stores buffers.

    **Emit folding.**  When every field emit stores in a buffer is either
    a copy of the cell's own main field or an integer constant (the usual
    "every buffer carries the state" emit), buffer slot ``p`` of a cell is
    a fixed function of its main word: ``(main & keep_p) | const_p``
    (the pack masks each field to its width, and a field's bits survive
    unpack and repack unchanged).  Absorb then reads a neighbour's buffer
    as that function of the neighbour's main word in LDS, and the kernel
    needs neither the emit pass nor the buffer planes.  ``fold`` holds
    ``{slot: (keep, const)}`` when this applies, else ``None``; it is
    decided by :meth:`start_absorb` after emit has been translated.
    Slots emit never stores read as zero, the reference's zero-initialised
    planes; under a static border the fold is used only if it maps the
    border word to itself in every slot (an off-board neighbour reads as
    the border word in every slot).  A buffer field emit leaves unassigned
    holds an indeterminate value in the reference; folding gives it 0.
    """

    name = "tile"

    def __init__(self, model):
        self._model = model
        self.phase = None
        self.emit_slots = set()
        self.emit_writes_main = False
        self.emit_reads_neighbor_main = False
        self.emit_reads_neighbor_buffer = False
        self.absorb_writes_buffer = False
        self.absorb_writes_main = False
        self.fold = None
        self.absorb_code = self.color_code = self.store_code = ""
        self.color_is_pure = False
        self.assign_count = 0
        self._assigned = {}
        self._slot_fold = {}

    def note_assignment(self, prop, value):
        """Count assignments; remember the last value of each during emit."""
        self.assign_count += 1
        if self.phase == "emit":
            self._assigned[prop.var_name] = str(value)

    def _fold_of(self, container):
        """``(keep, const)`` of a buffer word emit stores, or ``None``."""
        main = self._model.main
        keep = const = shift = 0
        for name, prop in container.properties.items():
            mask = (1 << prop.bit_width) - 1
            rhs = self._assigned.get(prop.var_name)
            if rhs is None:
                pass
            elif rhs == main[name].var_name:
                keep |= mask << shift
            elif _INT_LITERAL.match(rhs):
                const |= (int(rhs.strip("() ")) & mask) << shift
            else:
                return None
            shift += prop.bit_width
        return keep, const

    def start_absorb(self, fold=True):
        """End the emit phase: decide the fold, then translate absorb."""
        self.fold = self._decide_fold() if fold else None
        self.phase = "absorb"

    def _decide_fold(self):
        if self.emit_reads_neighbor_buffer or self.emit_writes_main:
            return None
        folds = {}
        for slot in range(1, len(self._model.buffers) + 1):
            if slot in self.emit_slots:
                fold = self._slot_fold.get(slot)
                if fold is None:
                    return None
            else:
                fold = (0, 0)
            folds[slot] = fold
        border = self._model.topology.border
        if hasattr(border, "off_board_state"):
            bits = self._model.main.best_type[0]
            value = int(border.value) & ((1 << bits) - 1)
            if any((value & keep) | const != value
                   for keep, const in folds.values()):
                return None
        return folds

    def _literal(self, value):
        """C expression of a word-sized constant."""
        if value < 1 << 64:
            return "((_word) %dULL)" % value
        return "((((_word) %dULL) << 64) | (_word) %dULL)" % (
            value >> 64, value & ((1 << 64) - 1))

    def _offset(self, nbr_num):
        delta = self._model.topology.neighborhood.neighbor_deltas[nbr_num]
        terms = []
        stride = ""
        for axis, step in enumerate(delta):
            if axis == 0:
                stride = ""
            elif axis == 1:
                stride = "_RW"
            else:
                stride = "_RW * _RH"
            if step:
                terms.append("(%d)%s" % (step, (" * " + stride) if stride else ""))
        return " + ".join(terms) if terms else "0"

    def _plane(self, buf_num, write=False):
        if buf_num == 0:
            return "_mp[_cur ^ 1]" if write else "_mp[_cur]"
        return "_tb + %d * _RA" % (buf_num - 1)

    def cell_ref(self, container):
        """Load expression of a container word."""
        index = "_li"
        if container.nbr_num >= 0:
            if self.phase == "emit":
                if container.buf_num == 0:
                    self.emit_reads_neighbor_main = True
                else:
                    self.emit_reads_neighbor_buffer = True
            index = "_li + %s" % self._offset(container.nbr_num)
        if (self.phase == "absorb" and self.fold is not None
                and container.buf_num > 0):
            keep, const = self.fold[container.buf_num]
            return "((_mp[_cur][%s] & %s) | %s)" % (
                index, self._literal(keep), self._literal(const))
        return "(%s)[%s]" % (self._plane(container.buf_num), index)

    def store_ref(self, container):
        """Store target of a container word."""
        if container.nbr_num >= 0:
            raise NotImplementedError("stores to neighbour words")
        if container.buf_num == 0:
            if self.phase == "emit":
                self.emit_writes_main = True
            elif self.phase == "absorb":
                self.absorb_writes_main = True
            return "(%s)[_li]" % self._plane(0, write=True)
        if self.phase == "emit":
            self.emit_slots.add(container.buf_num)
            self._slot_fold[container.buf_num] = self._fold_of(container)
        else:
            self.absorb_writes_buffer = True
        return "(%s)[_li]" % self._plane(container.buf_num)

    def neighbor_declaration(self, container):
        """Load of a neighbour word from the region."""
        if container.init_val is None:
            return "%s %s;\n" % (container.ctype, container.var_name)
        return "%s %s = %s;\n" % (container.ctype, container.var_name,
                                   self.cell_ref(container))
