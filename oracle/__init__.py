"""CPU FP64 oracle (test infrastructure only; see kt_oracle.cpp header)."""
