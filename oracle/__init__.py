"""CPU oracle (test infrastructure only): see oracle.py and uwg_oracle.c."""
