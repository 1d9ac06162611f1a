"""
The render epilogue's pixel tables (``runtime/fused.py:render_tables``)
against the plain renderer's per-pixel formula evaluated pixel by pixel
(``xentica/core/renderers.py:220-270``: ``(int) ((float) p / (float)
zoom) + shift``, wrapped onto the torus), and the kernel source of the
epilogue form compiled for gfx950.
"""
import struct

import numpy as np
import pytest

from xentica_amd.runtime.fused import render_tables


def f32(value):
    """Round to IEEE single precision."""
    return struct.unpack("f", struct.pack("f", value))[0]


def cell_of(pixel, zoom, shift, cells):
    """The renderer body's cell of one pixel coordinate (C semantics)."""
    x = int(f32(f32(float(pixel)) / f32(float(zoom)))) + shift
    if x < 0:
        x = cells - (-x % cells)
    if x >= cells:
        x = x % cells
    return x


@pytest.mark.parametrize("size,viewport,zoom,pos", [
    ((640, 360), (640, 360), 1, (0, 0)),
    ((640, 360), (200, 120), 3, (17, 5)),
    ((640, 360), (333, 77), 2, (630, 350)),
    ((640, 360), (1000, 500), 1, (600, 300)),
    ((100, 60), (700, 410), 7, (99, 59)),
    ((96, 64), (64, 48), 4, (5, 63)),
    ((37, 23), (300, 200), 1, (36, 0)),
])
def test_tables_list_every_pixel_under_its_cell(size, viewport, zoom, pos):
    width, height = viewport
    table = render_tables(size, width, height, zoom, *pos)
    w0, w1 = size
    assert table.dtype == np.int32
    assert table.size == (w0 + 1) + (w1 + 1) + width + height
    xoff, yoff = table[:w0 + 1], table[w0 + 1:w0 + w1 + 2]
    xpix = table[w0 + w1 + 2:w0 + w1 + 2 + width]
    ypix = table[w0 + w1 + 2 + width:]
    for offsets, pixels, count, shift, cells in (
            (xoff, xpix, width, pos[0], w0), (yoff, ypix, height, pos[1],
                                              w1)):
        assert offsets[0] == 0 and offsets[-1] == count
        assert np.all(np.diff(offsets) >= 0)
        assert sorted(pixels.tolist()) == list(range(count))
        for cell in range(cells):
            for pixel in pixels[offsets[cell]:offsets[cell + 1]]:
                assert cell_of(int(pixel), zoom, shift, cells) == cell


def test_tile_runs_cover_the_image_once():
    """Summed over a grid of tiles, each tile's run of columns times its
    run of rows lists every pixel exactly once."""
    size, (width, height), zoom, pos = (130, 70), (333, 211), 3, (120, 66)
    table = render_tables(size, width, height, zoom, *pos)
    w0, w1 = size
    xoff, yoff = table[:w0 + 1], table[w0 + 1:w0 + w1 + 2]
    xpix = table[w0 + w1 + 2:w0 + w1 + 2 + width]
    ypix = table[w0 + w1 + 2 + width:]
    seen = np.zeros(width * height, dtype=np.int64)
    tx, ty = 64, 16
    for oy in range(0, w1, ty):
        for ox in range(0, w0, tx):
            cols = xpix[xoff[ox]:xoff[min(ox + tx, w0)]]
            rows = ypix[yoff[oy]:yoff[min(oy + ty, w1)]]
            for py in rows:
                seen[py * width + cols] += 1
    assert np.all(seen == 1)


def test_epilogue_source_compiles():
    """The epilogue forms of Life (last-step colour form and the every-step
    one) and EvoLife (SMOOTH_FACTOR) build with hiprtc."""
    from examples.game_of_life import GameOfLife, GOLExperiment
    from examples.evolife import EvoLife, CrossbreedingExperiment
    from xentica_amd.runtime import hip
    from xentica_amd.runtime.fused import FusedProgram, _epilogue_body
    for cls, exp, forms in ((GameOfLife, GOLExperiment, 2),
                            (EvoLife, CrossbreedingExperiment, 1)):
        model = cls(exp)
        program = FusedProgram(model)
        body = _epilogue_body(model)
        assert body is not None and "_col[" in body
        assert len(program._forms) == forms
        for fast in program._forms:
            source = program._render(*program._forms[fast],
                                     render=(body, ""))
            assert "const unsigned i = " in source
            assert hip.compile_source(source, "epilogue.hip")
