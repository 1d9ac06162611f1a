"""
The reference's ConservedLife and PoissonWalk examples translate on this
package's DSL (CPU only): 4-D von Neumann topology, outer totalistic
rules with genome crossover, float ``exp`` in the rule, and emit storing
the main state, which keeps them on the plain engine.
"""
import pytest

from examples.conserved_life import (
    ConservedLife, ConservedLife4D, BigBangExperiment, BigBangExperiment4D)
from examples.poisson_walk import PoissonWalk, GlyphExperiment
from xentica_amd.core.exceptions import XenticaException
from xentica_amd.runtime.fused import FusedProgram


@pytest.mark.parametrize("cls,exp,k", [
    (ConservedLife, BigBangExperiment, 8),
    (ConservedLife4D, BigBangExperiment4D, 8),
    (PoissonWalk, GlyphExperiment, 8)])
def test_translates_with_plain_structure(cls, exp, k):
    model = cls(exp)
    assert len(model.buffers) == k
    src = model.cuda_source
    assert "void emit(" in src and "void absorb(" in src
    # 63 and 47 bits of state: one 64-bit container word
    assert model.main.ctype == "unsigned long"


def test_float_rule_and_emit_store():
    src = PoissonWalk(GlyphExperiment).cuda_source
    assert "exp(" in src


@pytest.mark.parametrize("cls,exp", [(ConservedLife, BigBangExperiment),
                                     (PoissonWalk, GlyphExperiment)])
def test_fused_engine_refuses_emit_storing_main(cls, exp):
    with pytest.raises(XenticaException, match="emit\\(\\) stores the main"):
        FusedProgram(cls(exp), tile=(32, 32), depth=2)
