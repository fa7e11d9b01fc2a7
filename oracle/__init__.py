"""CPU oracle for the local SpGEMM path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this package, and only as the checker.  The product path never does.
"""
from .oracle import *  # noqa: F401,F403
