"""CPU FP64 oracle of the aceeric/nbodygo per-cycle compute path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package,
and only as the checker (or, for the CPU baseline, as the timed reference algorithm).  The product
never routes through it.
"""
