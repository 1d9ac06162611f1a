"""Example models and experiments, written in the xentica_amd DSL."""
