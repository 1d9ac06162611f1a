"""Seeds: random streams and patterns producing a model's step-0 state."""
from xentica_amd.seeds import random, patterns  # noqa: F401

__all__ = ['random', 'patterns', ]
