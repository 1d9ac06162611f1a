"""CPU emulations of GPU kernels (test infrastructure)."""
