"""
The reference's import name for this engine.

User code written for Xentica imports ``xentica`` (``from xentica import
core``, ``from xentica.core import CellularAutomaton``, ``from
xentica.seeds.random import RandInt``, ``from xentica.tools.rules import
LifeLike``, ...; ``xentica/core/__init__.py`` and the examples of the
reference).  Every ``xentica.<name>`` module is the module
``xentica_amd.<name>`` itself, registered under both names, so classes
and parameters are the same objects whichever name a module was imported
by.
"""
import importlib
import importlib.abc
import importlib.util
import sys

from xentica_amd import __version__  # noqa: F401

_PREFIX = __name__ + "."
_TARGET = "xentica_amd."


class _AliasFinder(importlib.abc.MetaPathFinder, importlib.abc.Loader):
    """Resolves ``xentica.<name>`` to the module ``xentica_amd.<name>``."""

    def find_spec(self, fullname, path=None, target=None):
        """A spec for ``xentica.<name>`` when ``xentica_amd.<name>`` exists."""
        if not fullname.startswith(_PREFIX):
            return None
        real = _TARGET + fullname[len(_PREFIX):]
        try:
            found = importlib.util.find_spec(real)
        except ModuleNotFoundError:
            return None
        if found is None:
            return None
        return importlib.util.spec_from_loader(
            fullname, self, is_package=found.submodule_search_locations
            is not None)

    def create_module(self, spec):
        """The ``xentica_amd`` module itself."""
        return importlib.import_module(_TARGET + spec.name[len(_PREFIX):])

    def exec_module(self, module):
        """Nothing to run: the module was executed under its own name."""


if not any(isinstance(finder, _AliasFinder) for finder in sys.meta_path):
    sys.meta_path.insert(0, _AliasFinder())
