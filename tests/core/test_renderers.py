"""Renderers (``tests/core/test_renderers.py`` of the reference)."""
from xentica_amd.core.renderers import Renderer, RendererPlain


def test_base():
    assert Renderer().render_code() == ""


def test_renderer_plain():
    assert RendererPlain(projection_axes=(1, 0)).projection_axes == (1, 0)
