"""Test infrastructure: the CPU oracle (a C restatement of the reference step,
oracle/b2oracle.c) and the recipe that builds the reference itself (oracle/Makefile).
Only tests, the smoke check and bench.py's CPU-baseline leg use this package."""
