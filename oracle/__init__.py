"""CPU oracle of the per-step granular pipeline -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
as the checker and as the timed CPU baseline; the product never does.
"""
