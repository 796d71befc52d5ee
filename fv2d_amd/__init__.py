"""fv2d_amd: MI355X-native explicit 2D compressible-hydro step (fv2d hot path).

The compute path is libfv2d_amd.so (hand-written HIP kernels for gfx950 behind the C ABI
of include/fv2d_amd.h). `Solver` is the Python host mirror of that ABI.
"""
from .capi import Params  # noqa: F401
