"""FP64 oracle (test infrastructure; never imported by the product package)."""
