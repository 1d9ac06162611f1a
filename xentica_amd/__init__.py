"""
xentica_amd: a from-scratch Xentica engine for AMD MI355X (gfx950).

The modelling surface of Xentica (``core``, ``seeds``, ``tools``,
``bridge``, ``utils``) is kept as a drop-in; underneath, models are
translated to HIP C++ and run through ``libxentica_hip.so``.
"""
__version__ = "0.1.0"
