"""ORACLE — test infrastructure only: CPU restatements of the reference
ForceAtlas2 algorithms (fa2_oracle.c via ctypes, ref_numpy.py). Only tests/,
the smoke check and the CPU baseline of bench.py use it."""
