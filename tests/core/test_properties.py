"""Properties (``tests/core/test_properties.py`` of the reference), CPU part."""
import pytest

from xentica_amd.core.properties import (
    Property, IntegerProperty, ContainerProperty, TotalisticRuleProperty,
)
from xentica_amd.core.exceptions import XenticaException
from examples.game_of_life import GameOfLife, GOLExperiment


def test_width():
    assert IntegerProperty(max_val=1).width == 1


def test_default_bit_width():
    assert Property().bit_width == 1


def test_broad_bit_width():
    prop = IntegerProperty(max_val=10e23)
    assert prop.bit_width == 80
    assert prop.width == 2


def test_set():
    prop = Property()
    prop.cont = ContainerProperty()
    prop.cont.var_name = "foo"
    prop.cont = 1
    assert type(prop.cont).__name__ == "ContainerProperty"


def test_unbound():
    prop = Property()
    with pytest.raises(XenticaException):
        assert not prop.declared
    with pytest.raises(XenticaException):
        assert prop.coords_declared
    with pytest.raises(XenticaException):
        cprop = ContainerProperty()
        assert not cprop.unpacked


def test_container_values():
    assert len(list(GameOfLife.main.values())) == 1


def test_pure_totalistic():
    prop = TotalisticRuleProperty(outer=False)
    with pytest.raises(XenticaException):
        prop.is_sustained(8)
    with pytest.raises(XenticaException):
        prop.is_born(8)
    prop.var_name = "rule"
    assert str(prop.next_val(1, 3)) == "((rule >> 4) & 1)"


def test_outer_totalistic():
    class GolTest(GameOfLife):
        rule = TotalisticRuleProperty(outer=True)
    model = GolTest(GOLExperiment)
    assert str(model.main.rule.is_sustained(8)) == \
        "(((_cell_rule & 261120) >> 17) & 1)"
    assert str(model.main.rule.is_born(8)) == "(((_cell_rule & 510) >> 8) & 1)"
    assert str(model.main.rule.next_val(1, 3)) == "((_cell_rule >> 12) & 1)"


def test_bit_widths_and_container_type():
    """Container word type from the summed widths (``properties.py:70-151``)."""
    class Wide(GameOfLife):
        a = IntegerProperty(max_val=255 * 2)
        rule = TotalisticRuleProperty(outer=True)
    model = Wide(GOLExperiment)
    assert model.main.a.bit_width == 9
    assert model.main.rule.bit_width == 18
    assert model.main.bit_width == 1 + 9 + 18
    assert model.main.ctype == "unsigned int"
