"""FP64 CPU oracle for the DP Stokes Mdot path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package, and only as the checker.
"""
