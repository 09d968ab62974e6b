"""CPU checkers for the FLIP transfer + advection path (test infrastructure only)."""
