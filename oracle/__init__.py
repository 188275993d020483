"""CPU oracle of the SCBA inner loop (test infrastructure only)."""
