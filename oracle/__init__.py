"""CPU oracle for the cell-tree hot path (test infrastructure only; never imported by the product)."""
