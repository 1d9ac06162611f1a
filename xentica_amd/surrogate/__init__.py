"""
Neural-CA surrogate of a coarse observable (no counterpart in the reference).

``network`` -- the hand-written gfx950 kernels behind a
``torch.autograd.Function`` and the :class:`NCASurrogate` module;
``train`` -- training pairs from the exact engine, training, evaluation
against the exact engine and the persistence baseline.
"""
