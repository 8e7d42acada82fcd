"""FP64 CPU oracle for the QJX dimer trajectory path (test infrastructure only)."""
