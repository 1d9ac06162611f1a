"""
Topology of a model: dimensionality, lattice, neighbourhood and border.

Every ``CellularAutomaton`` subclass declares an inner ``Topology`` class
with ``dimensions``, ``lattice``, ``neighborhood`` and ``border``
(``xentica/core/topology/__init__.py``)::

    class Topology:
        dimensions = 2
        lattice = core.OrthogonalLattice()
        neighborhood = core.MooreNeighborhood()
        border = core.TorusBorder()
"""
