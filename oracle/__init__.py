"""CPU oracle for the XR hot path (test infrastructure only; see oracle.py)."""
