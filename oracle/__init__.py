"""CPU oracle for the ParFlex hot path (test infrastructure only)."""
