"""
Renderers: project cell colours onto a 2-D image.

:class:`RendererPlain` (``xentica/core/renderers.py:115-270``) maps each
pixel through zoom and pan to a cell on the two projection axes, wraps
toroidally, averages the colours over all remaining axes and divides by
``SMOOTH_FACTOR`` when the model defines it.  Key actions scroll and zoom
(``renderers.py:156-218``).

Kernel arguments are ``(int3 *col, int *img, int zoom, int dx, int dy,
int width)`` followed by the pixel count.
"""
import abc

import numpy as np

from xentica_amd.core.mixins import BscaDetectorMixin

__all__ = ['Renderer', 'RendererPlain', ]


class Renderer(BscaDetectorMixin):
    """Base class for all renderers."""

    def __init__(self):
        """Initialize kernel arguments."""
        self.args = [
            ("int3", "*col"),
            ("int", "*img"),
        ]

    def get_args_vals(self, bsca):
        """Argument values, in the order of ``self.args``."""
        return [bsca.gpu.arrays.colors, bsca.gpu.arrays.img]

    @abc.abstractmethod
    def render_code(self):
        """C code of the per-pixel render body."""
        return ""

    def setup_actions(self, bridge):
        """Add renderer-specific key actions to ``bridge``."""


class RendererPlain(Renderer):
    """Render the board as a 2-D plane, projecting other axes by averaging."""

    def __init__(self, projection_axes=None):
        """Initialize kernel arguments and projection axes."""
        super().__init__()
        self.args += [
            ("int", "zoom"),
            ("int", "dx"),
            ("int", "dy"),
            ("int", "width"),
        ]
        self.projection_axes = projection_axes
        if self.projection_axes is None:
            self.projection_axes = (0, 1)

    def get_args_vals(self, bsca):
        """Extend the argument values with zoom, pan and width."""
        args_vals = super().get_args_vals(bsca)
        args_vals += [
            np.int32(bsca.zoom),
            np.int32(bsca.pos[0]),
            np.int32(bsca.pos[1]),
            np.int32(bsca.width),
        ]
        return args_vals

    def setup_actions(self, bridge):
        """Scroll with the arrow keys, zoom with ``=`` and ``-``."""
        bridge.key_actions.update({
            "up": self.move(0, 1),
            "down": self.move(0, -1),
            "left": self.move(-1, 0),
            "right": self.move(1, 0),
            "=": self.zoom(1),
            "-": self.zoom(-1),
        })

    @staticmethod
    def move(delta_x, delta_y):
        """Action moving the viewport by a delta."""
        def func(model, _gui):
            model.renderer.apply_move(model, delta_x, delta_y)
        return func

    @staticmethod
    def zoom(dzoom):
        """Action changing the zoom by a delta."""
        def func(model, _gui):
            model.renderer.apply_zoom(model, dzoom)
        return func

    @staticmethod
    def apply_move(bsca, *args):
        """Move ``bsca.pos`` toroidally."""
        for axis, delta in enumerate(args):
            bsca.pos[axis] = (bsca.pos[axis] + delta) % bsca.size[axis]

    @staticmethod
    def apply_zoom(bsca, dval):
        """Change ``bsca.zoom``, never below 1."""
        bsca.zoom = max(1, (bsca.zoom + dval))

    def render_code(self):
        """C code of the per-pixel body (``renderers.py:220-270``)."""
        bsca = self.bsca
        width = bsca.topology.lattice.width_prefix
        ax0, ax1 = self.projection_axes
        xname, yname = "x%d" % ax0, "x%d" % ax1
        code = (
            "int {x} = (int) (((float) (i % width)) / (float) zoom) + dx;\n"
            "int {y} = (int) (((float) (i / width)) / (float) zoom) + dy;\n"
            "if ({x} < 0) {x} = {w}{i0} - (-{x} % {w}{i0});\n"
            "if ({x} >= {w}{i0}) {x} = {x} % {w}{i0};\n"
            "if ({y} < 0) {y} = {w}{i1} - (-{y} % {w}{i1});\n"
            "if ({y} >= {w}{i1}) {y} = {y} % {w}{i1};\n"
            "float r = 0, g = 0, b = 0;\n"
            "int num_cells_projected = 1;\n"
        ).format(x=xname, y=yname, w=width, i0=ax0, i1=ax1)
        loops = ""
        for axis in range(bsca.topology.dimensions):
            if axis in self.projection_axes:
                continue
            code += "num_cells_projected *= {w}{i};\n".format(w=width, i=axis)
            loops += "for (int x{i} = 0; x{i} < {w}{i}; x{i}++) {{\n".format(
                w=width, i=axis)
        code += loops
        code += (
            "int ii = {idx};\n"
            "int3 c = col[ii];\n"
            "r += c.x;\n"
            "g += c.y;\n"
            "b += c.z;\n"
        ).format(idx=bsca.topology.lattice.coord_to_index_code("x"))
        code += "}" * (bsca.topology.dimensions - 2)
        smooth = ""
        if bsca.is_constant("SMOOTH_FACTOR"):
            smooth = "/ SMOOTH_FACTOR"
        code += (
            "img[i * 3] = r / num_cells_projected {s};\n"
            "img[i * 3 + 1] = g / num_cells_projected {s};\n"
            "img[i * 3 + 2] = b / num_cells_projected {s};\n"
        ).format(s=smooth)
        return code
