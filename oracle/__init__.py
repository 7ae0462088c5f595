"""CPU FP64 oracle for the edge-state Monte-Carlo path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this package,
and only as the checker.  The product path (laughlinedge_amd) never imports it.
"""
