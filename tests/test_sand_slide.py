"""
The sliding-sand rule (``examples/sand_slide.py``): its generated code
(rendered for the host) equals a NumPy restatement of the rule step by
step, conserves the total height and keeps heights within 0-8.
"""
import numpy as np

from examples.sand_slide import SandSlide, slide_experiment
from xentica_amd.runtime.hostc import PlainHost


def shift(a, dy, dx):
    return np.roll(np.roll(a, dy, 0), dx, 1)


def numpy_step(grains, odd):
    if odd:
        return shift(grains, 1, 1)
    out = grains.copy()
    for dy, dx in ((0, 1), (1, 0)):
        slide = (grains >= shift(grains, -dy, -dx) + 3).astype(grains.dtype)
        out = out - slide + shift(slide, dy, dx)
    return out


def test_generated_code_equals_numpy_restatement():
    model = SandSlide(slide_experiment(5, (96, 80)))
    host = PlainHost(model)
    host.load(model.gpu.arrays.cells.get(), model.gpu.arrays.colors.get())
    grains = (host.cells & 15).astype(np.int64).reshape(80, 96)
    total = grains.sum()
    for step in range(48):
        host.step()
        grains = numpy_step(grains, step % 2 == 1)
        cells = host.cells.reshape(80, 96)
        assert np.array_equal(cells & 15, grains), step
        assert np.all(cells >> 4 == (step + 1) % 2), step
        assert grains.sum() == total
        assert grains.min() >= 0 and grains.max() <= 8


def test_slides_happen_and_stop_being_a_plain_roll():
    """The first 16 steps differ from the pure half-block drift (slides
    moved grains); the drift alone is the roll by (8, 8)."""
    rng = np.random.default_rng(0)
    grains = rng.integers(0, 9, (64, 64))
    start = grains.copy()
    for step in range(16):
        grains = numpy_step(grains, step % 2 == 1)
    assert not np.array_equal(grains, shift(start, 8, 8))
    assert grains.sum() == start.sum()
