"""CPU oracle for the ncarray device hot path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this package,
as the checker.  The product (ncarray_amd) never imports it.
"""
