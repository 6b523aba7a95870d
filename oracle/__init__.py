"""Test infrastructure: the CPU oracle of the sieve search path (see oracle/oracle.py)."""
