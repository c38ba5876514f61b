"""TEST INFRASTRUCTURE ONLY: the FP64 CPU oracle of the BEEP hot path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, as the checker; the product (beep_amd/) never does.

Parity status per function is listed in oracle/beep_oracle.cc's header."""
