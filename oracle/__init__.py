"""CPU oracle of the pressure-solve path -- test infrastructure only (see oracle.py)."""
