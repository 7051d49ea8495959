"""FP64 oracle of the per-frame filtering path -- test infrastructure only (see oracle.py)."""
