"""
Which model is being translated right now.

The reference finds the model that DSL objects emit code into by walking
interpreter frames for a ``CellularAutomaton`` instance
(``xentica/core/mixins.py:23-40``).  Here the model under translation is
pushed on an explicit per-thread stack by the translator itself, so code
always lands in the right model even when several models translate in
one process.  Outside a translation (DSL objects poked at from user code
or tests after construction) the frame walk is kept as the fallback, so
the observable behaviour of the reference is preserved.
"""
import threading

_STATE = threading.local()


def _stack():
    stack = getattr(_STATE, "stack", None)
    if stack is None:
        stack = []
        _STATE.stack = stack
    return stack


class translating:  # pylint: disable=invalid-name
    """Context manager making ``model`` the target of emitted code."""

    def __init__(self, model):
        self._model = model

    def __enter__(self):
        _stack().append(self._model)
        return self._model

    def __exit__(self, *exc):
        _stack().pop()
        return False


def current():
    """Return the model under translation, or ``None``."""
    stack = _stack()
    return stack[-1] if stack else None
