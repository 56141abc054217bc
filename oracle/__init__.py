"""CPU oracle of the cloth substep (test infrastructure only; see xrt_oracle.h)."""
