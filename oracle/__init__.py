"""CPU checkers for the SpMV path -- test infrastructure only (see oracle.py)."""
