"""FP64 CPU oracle of pantea's NNP path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package.
"""
