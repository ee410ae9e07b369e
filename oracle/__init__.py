"""FP64 CPU oracle of the acoustic_petsc numerics (test infrastructure only)."""
