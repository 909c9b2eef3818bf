"""CPU f64 oracle for the BlochHighOrder path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this package.  It is the checker, never the thing measured as the
product or shipped: the product path (highordermrisim.jl_amd) never imports it.

Parity status: every function is pinned by a closed-form or brute-force check
in tests/test_oracle_kats.py (see the header of hom_oracle.c for the list).
"""
from .oracle import (available, build, block_adc, combine, excitation, partition,  # noqa: F401
                     precession, run, run_workload)
