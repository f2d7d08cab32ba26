"""FP64 CPU oracle of the reference SPH solver path (test infrastructure)."""
