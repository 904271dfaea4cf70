"""CPU oracle of the reference arithmetic: test infrastructure only (see lf_oracle.h)."""
