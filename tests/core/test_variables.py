"""DSL variables (``tests/core/test_variables.py`` of the reference)."""
import pytest

from xentica_amd.core.variables import Variable, IntegerVariable
from xentica_amd.core.exceptions import XenticaException
from xentica_amd import core
from xentica_amd import seeds


class InvertCA(core.CellularAutomaton):
    """CA that assigns through a class-level variable."""

    state = core.IntegerProperty(max_val=1)
    intvar = IntegerVariable()

    class Topology:
        dimensions = 2
        lattice = core.OrthogonalLattice()
        neighborhood = core.MooreNeighborhood()
        border = core.TorusBorder()

    def emit(self):
        self.intvar = self.main.state
        self.buffers[0].state = self.intvar

    def absorb(self):
        self.intvar = self.buffers[0].state
        self.main.state = -self.intvar
        var_list = [
            IntegerVariable(name="self_var"),
            IntegerVariable(),
        ]
        var_list[1] += 1

    def color(self):
        """No colour processing."""


class InvertExperiment(core.Experiment):
    word = "INSIDE OUT"
    size = (64, 36, )
    seed = seeds.patterns.PrimordialSoup(
        vals={"state": seeds.random.RandInt(0, 1)})


def test_integer():
    """Names follow the caller's locals (``test_variables.py:57-78``)."""
    myvar1 = IntegerVariable()
    assert str(myvar1) == "myvar1"
    assert myvar1.code == "myvar1"
    myvar2 = IntegerVariable(1)
    assert str(myvar2) == "myvar2"
    var_list = [
        IntegerVariable(1, name="myvar3"),
        IntegerVariable(name="myvar4"),
        IntegerVariable(2),
        IntegerVariable(),
    ]
    assert var_list[0].var_name == "myvar3"
    assert var_list[1].var_name == "myvar4"
    assert var_list[2].var_name == "var"
    assert var_list[3].var_name == "var"


def test_descriptor():
    """Exact generated text (``test_variables.py:80-115``)."""
    model = InvertCA(InvertExperiment)
    correct_emit = """
unsigned char _cell_state;
int tmp__cell;
tmp__cell = (_cell) & 1;
_cell_state = tmp__cell;
unsigned int intvar = 0;
intvar = _cell_state;
unsigned char _bcell_state0;
_bcell_state0 = intvar;
unsigned char _bcell0;
_bcell0 = ((unsigned char) _bcell_state0 & 1);
fld[i + n * 8] = _bcell0;
"""
    correct_absorb = """
unsigned char _bcell_state0;
int tmp__bcell0;
tmp__bcell0 = (_bcell0) & 1;
_bcell_state0 = tmp__bcell0;
unsigned int intvar = 0;
intvar = _bcell_state0;
unsigned char _cell_state;
_cell_state = (-(intvar));
unsigned char _cell = fld[i];
unsigned int var = 0;
var += 1;
_cell = ((unsigned char) _cell_state & 1);
fld[i] = _cell;
"""
    assert correct_emit in model.cuda_source
    assert correct_absorb in model.cuda_source


def test_illegal_assign():
    """``test_variables.py:117-129``."""
    with pytest.raises(XenticaException):
        class BrokenCA(InvertCA):
            def emit(self):
                deferred_exp = 1 + self.main.state
                deferred_exp += 1
        BrokenCA(InvertExperiment)


def test_no_init_val():
    """``test_variables.py:131-134``."""
    with pytest.raises(XenticaException):
        Variable()
