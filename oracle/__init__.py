"""ORACLE — test infrastructure only (the checker, never the product path)."""
