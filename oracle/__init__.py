"""CPU oracle for the ducc_amd NUFFT (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this package, and only as the checker.  The product (ducc_amd) never imports
it.  `oracle.core` wraps the C restatement in oracle.c; `oracle.ref` wraps the
compiled reference ducc0 in oracle/_ref (present only where it was built).
"""
