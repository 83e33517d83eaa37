"""CPU oracle for the hybrid-accumulation SpGEMM (test infrastructure only)."""
