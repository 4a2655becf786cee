"""CPU oracle package (test infrastructure only; see oracle/kpc_oracle.h)."""
