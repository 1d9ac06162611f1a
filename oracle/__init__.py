"""CPU oracle of the reference's semantics -- test infrastructure only."""
