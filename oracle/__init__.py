"""CPU oracle of the back-projection -- test infrastructure (see oracle/sarbp_oracle.c)."""
