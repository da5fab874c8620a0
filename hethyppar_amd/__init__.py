"""MI355X-native BiCGStab / SpMV engine: one rank's share of HetHypPar's solve.

The compute path is libhbs.so (hand-written HIP kernels for gfx950 behind a
C-ABI, include/hbs.h); `hethyppar_amd.hbs` is its ctypes binding.
"""
from . import hbs  # noqa: F401

__all__ = ["hbs"]
