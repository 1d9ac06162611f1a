"""
Multi-property, RNG and float models of the reference on the plain engine.

Restates ``test_genetics_general`` (``tests/core/test_base.py:197-216``) and
``test_float`` (``tests/core/test_properties.py:97-104``) with their golden
CRC-32s.  The two EvoLife runs stay in one test, in the reference's order:
``Parameter`` objects are class-level, so the second experiment's values
persist (``xentica/core/parameters.py:143-145``).
"""
import binascii

import pytest

from examples.evolife import (
    EvoLife, CrossbreedingExperiment, CrossbreedingExperiment2)
from examples.gravicowboys import GraviCowboys, GraviCowboysExperiment

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["plain", "auto"], autouse=True)
def _engine(request, monkeypatch):
    """Run every model on the plain engine and on the default one."""
    from xentica_amd.core.base import CellularAutomaton
    monkeypatch.setattr(CellularAutomaton, "engine", request.param)


def run(model, steps):
    for _ in range(steps):
        model.step()
    return model


def crcs(model):
    cells = model.gpu.arrays.cells.get()[:model.cells_num]
    return binascii.crc32(cells), binascii.crc32(model.gpu.arrays.colors.get())


def test_genetics_general():
    # parameters are class-level (xentica/core/parameters.py:143-145): start
    # from the defaults, as the reference's test does in a fresh process
    for name, value in (("death_speed", 15), ("max_genes", 9),
                        ("mutation_prob", 0.0)):
        EvoLife.__dict__[name].__set__(None, value)
    model = run(EvoLife(CrossbreedingExperiment, legacy_coloring=False), 100)
    cells, colors = crcs(model)
    assert cells == 2061779088
    assert colors == 68940598
    model = run(EvoLife(CrossbreedingExperiment2), 100)
    cells, colors = crcs(model)
    assert cells == 3381214984
    assert colors == 2253830262


def test_float():
    model = run(GraviCowboys(GraviCowboysExperiment), 100)
    assert crcs(model)[0] == 1068195512
