"""GUI bridges (``xentica/bridge``): key maps from a GUI to model actions."""
from xentica_amd.bridge.moire import MoireBridge

__all__ = ["MoireBridge", ]
