"""DSL helpers: math intrinsics, rule encodings, genetics and colour tools."""
from .xmath import Xmath as xmath

__all__ = ["xmath", ]
