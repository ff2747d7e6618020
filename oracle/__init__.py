"""CPU oracle (test infrastructure only; see cf_oracle.py)."""
