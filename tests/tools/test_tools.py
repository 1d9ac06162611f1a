"""Tools (``tests/tools/test_*.py`` of the reference)."""
from xentica_amd.tools.rules import LifeLike
from xentica_amd.tools import xmath


def test_golly():
    assert LifeLike.golly2int("B3/S23") == 6152
    assert LifeLike.int2golly(6152) == "B3/S23"
    assert LifeLike.int2golly(LifeLike.golly2int("B35678/S5678")) == \
        "B35678/S5678"


def test_xmath():
    assert str(xmath.popc("x")) == "(__popc(x))"
    assert str(xmath.exp("x")) == "(exp(x))"
    assert str(xmath.min("a", "b", "c")) == \
        "(((((a) < (b)) ? (a) : (b)) < (c)) ? (((a) < (b)) ? (a) : (b)) : (c))"
    assert str(xmath.int("x")) == "((unsigned int) (x))"
    assert str(xmath.float("x")) == "((float) (x))"
