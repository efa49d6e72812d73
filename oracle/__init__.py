"""CPU oracle of the DSMC step (test infrastructure only; see dsmc_oracle.h)."""
