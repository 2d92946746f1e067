"""CPU restatement of hera's EM -- test infrastructure (see oracle.py)."""
