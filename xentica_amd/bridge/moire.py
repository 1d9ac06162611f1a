"""Key map of the Moire GUI (``xentica/bridge/moire.py:28-34``)."""
from xentica_amd.bridge.base import Bridge

__all__ = ['MoireBridge', ]


class MoireBridge:
    """``[``/``]`` speed, space pauses, F12 system info, Esc exits."""

    key_actions = {
        "[": Bridge.speed(-1),
        "]": Bridge.speed(1),
        "spacebar": Bridge.toggle_pause,
        "f12": Bridge.toggle_sysinfo,
        "escape": Bridge.exit_app,
    }
