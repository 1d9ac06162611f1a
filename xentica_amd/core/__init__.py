"""
The modelling DSL: base class, properties, variables, parameters,
topology and experiments (``xentica/core/__init__.py``).

Shortcuts: ``core.CellularAutomaton``, ``core.Experiment``,
``core.OrthogonalLattice``, ``core.MooreNeighborhood``,
``core.VonNeumannNeighborhood``, ``core.TorusBorder``,
``core.StaticBorder``, ``core.IntegerProperty``, ``core.FloatProperty``,
``core.TotalisticRuleProperty``, ``core.RandomProperty``,
``core.XorShiftProperty`` (no reference counterpart),
``core.Parameter``, ``core.IntegerVariable``, ``core.FloatVariable``.
"""
from xentica_amd.core.base import CellularAutomaton
from xentica_amd.core.properties import (
    IntegerProperty,
    FloatProperty,
    TotalisticRuleProperty,
    RandomProperty,
    XorShiftProperty,
)
from xentica_amd.core.variables import (
    IntegerVariable, FloatVariable,
)
from xentica_amd.core.parameters import (
    Parameter,
)
from xentica_amd.core.topology.lattice import (
    OrthogonalLattice,
)
from xentica_amd.core.topology.neighborhood import (
    MooreNeighborhood, VonNeumannNeighborhood
)
from xentica_amd.core.topology.border import (
    TorusBorder, StaticBorder,
)
from xentica_amd.core.experiment import Experiment

__all__ = [
    'CellularAutomaton',
    'IntegerProperty',
    'FloatProperty',
    'TotalisticRuleProperty',
    'RandomProperty',
    'XorShiftProperty',
    'Parameter',
    'IntegerVariable',
    'FloatVariable',
    'OrthogonalLattice',
    'MooreNeighborhood',
    'VonNeumannNeighborhood',
    'TorusBorder',
    'StaticBorder',
    'Experiment',
]
