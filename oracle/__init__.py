"""Test-only CPU oracle for NRG's chemistry step.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package; it is the checker, never the thing shipped or measured as the
product.  See nrg_oracle.h for what each function restates and what pins it.
"""
from .pyoracle import Oracle, RefSlatec, build_oracle, ORACLE_SO, REF_SO  # noqa: F401
