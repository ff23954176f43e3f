"""MI355X-native pressure-solve hot path of freeCappuccino.

The product is ``libfcap.so`` (HIP kernels for gfx950 behind a C ABI declared
in include/fcap.h).  This package holds its ctypes front end, the host-side
mirror of the reference's solver interface, and the synthetic mesh/workload
generators.
"""
