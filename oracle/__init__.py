"""CPU oracle of the Greenstein-Winslow hot path (test infrastructure only)."""
