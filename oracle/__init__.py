"""CPU oracle of the HiDEM hot path -- test infrastructure only (see hdm_oracle.h)."""
