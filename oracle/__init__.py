"""CPU oracle of the D3Q27 cumulant hot path: test infrastructure only (see oracle/oracle.py)."""
