"""CPU oracle of the compressible RK3 step (test infrastructure only)."""
