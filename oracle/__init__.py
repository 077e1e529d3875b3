"""CPU oracle for the tau00 path -- test infrastructure only (see oracle.py)."""
