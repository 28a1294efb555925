"""CPU oracle (test infrastructure only): see oracle/sem_oracle.py."""
