"""CPU FP64 oracle (test infrastructure only; never imported by the product)."""
