"""Bridge (``tests/bridge/test_base.py`` of the reference)."""
import pytest

from xentica_amd.bridge.base import Bridge
from examples.game_of_life import GameOfLife, GOLExperiment


class SysInfo:
    def toggle(self):
        """Toggle."""


class DummyGui:
    def __init__(self):
        self.sysinfo = SysInfo()

    def exit_app(self):
        """Exit."""


@pytest.fixture
def model():
    return GameOfLife(GOLExperiment)


def test_noop_exit_sysinfo(model):
    gui = DummyGui()
    Bridge.noop(model, gui)
    Bridge.exit_app(model, gui)
    Bridge.toggle_sysinfo(model, gui)


@pytest.mark.parametrize("key,axis,delta", [
    ("up", 1, 1), ("down", 1, -1), ("right", 0, 1), ("left", 0, -1)])
def test_moves(model, key, axis, delta):
    coord = model.pos[axis]
    model.bridge.key_actions[key](model, DummyGui())
    assert model.pos[axis] == (coord + delta) % model.size[axis]


def test_zoom(model):
    zoom = model.zoom
    model.bridge.key_actions['='](model, DummyGui())
    assert model.zoom == zoom + 1
    model.bridge.key_actions['-'](model, DummyGui())
    model.bridge.key_actions['-'](model, DummyGui())
    assert model.zoom == max(zoom - 1, 1)


def test_speed_and_pause(model):
    speed = model.speed
    model.bridge.key_actions[']'](model, DummyGui())
    assert model.speed == speed + 1
    model.bridge.key_actions['['](model, DummyGui())
    model.bridge.key_actions['['](model, DummyGui())
    assert model.speed == max(speed - 1, 1)
    paused = model.paused
    Bridge.toggle_pause(model, DummyGui())
    assert model.paused == (not paused)
