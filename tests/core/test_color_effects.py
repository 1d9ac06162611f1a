"""
``MovingAverage`` (``xentica/core/color_effects.py:106-130``): the clamp
``max(min(new, old + FADE_IN), old - FADE_OUT)`` is emitted as the median
of ``(new, old - FADE_OUT, old + FADE_IN)`` when the bounds cannot cross,
and in the reference's form otherwise (CPU only).
"""
import numpy as np

from examples.game_of_life import GameOfLifeColor, GOLExperimentColor

MED3 = "max(min(new_r, old_col.x - FADE_OUT), min(max(new_r, old_col.x - FADE_OUT), old_col.x + FADE_IN))"
PLAIN = "max(min(new_r, old_col.x + FADE_IN), old_col.x - FADE_OUT)"


def source_with(fade_in, fade_out):
    exp = type("Fade%d_%d" % (fade_in, fade_out), (GOLExperimentColor, ),
               {"fade_in": fade_in, "fade_out": fade_out})
    return GameOfLifeColor(exp).cuda_source


def test_median_form_when_bounds_ordered():
    for fade_in, fade_out in ((255, 255), (18, 0), (0, 0), (-3, 5)):
        src = source_with(fade_in, fade_out)
        assert MED3 in src and PLAIN not in src


def test_reference_form_when_bounds_cross():
    src = source_with(-10, 3)
    assert PLAIN in src and MED3 not in src


def test_median_equals_clamp_for_ordered_bounds():
    """max(min(v, lo), min(max(v, lo), hi)) == max(min(v, hi), lo) for
    lo <= hi, over random and boundary values."""
    rng = np.random.default_rng(5)
    old = rng.integers(-1000, 1000, 20000)
    new = rng.integers(-2000, 2000, 20000)
    fade_in = rng.integers(-300, 300, 20000)
    fade_out = rng.integers(-300, 300, 20000)
    keep = fade_in + fade_out >= 0
    old, new, fade_in, fade_out = old[keep], new[keep], fade_in[keep], fade_out[keep]
    new[:50] = (old - fade_out)[:50]
    new[50:100] = (old + fade_in)[50:100]
    lo, hi = old - fade_out, old + fade_in
    med = np.maximum(np.minimum(new, lo), np.minimum(np.maximum(new, lo), hi))
    ref = np.maximum(np.minimum(new, hi), lo)
    assert np.array_equal(med, ref)
    # crossed bounds: the two forms differ, hence the guard
    lo, hi = np.array([5]), np.array([2])
    v = np.array([0])
    assert (np.maximum(np.minimum(v, lo), np.minimum(np.maximum(v, lo), hi))
            != np.maximum(np.minimum(v, hi), lo)).all()
