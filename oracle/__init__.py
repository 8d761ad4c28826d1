"""CPU oracle of the MPDATA path (test infrastructure only)."""
