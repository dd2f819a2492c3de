"""CPU oracle of the EZLDA sweep -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this package, as the checker or the timed CPU baseline; the product never
does.  See ezlda_oracle.c for the semantics and reference citations.
"""
