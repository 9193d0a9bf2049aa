"""CPU restatement of the futspace reference (test infrastructure only)."""
