"""CPU oracle (test infrastructure): C++ restatements of the reference twins
and FP64 definitions, loaded through oracle/oracle.py."""
