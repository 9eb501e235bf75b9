"""CPU oracle of the stereo hot path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use
this package. It wraps liboracle.so (tdvref.c) with numpy via ctypes.
"""
from .pyoracle import *  # noqa: F401,F403
