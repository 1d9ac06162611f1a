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

__all__ = ['PlainAddressing', ]


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
