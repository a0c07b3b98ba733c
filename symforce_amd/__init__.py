"""MI355X-native implementation of SymForce's optimizer hot path.

The C ABI lives in include/symhip.h (libsymhip.so, built in-tree by symforce_amd.build); the C++
host surface mirroring sym::Optimizer / sym::Values / sym::Factor lives in include/sym/. This
Python package is a thin ctypes binding used by the tests and the benchmark.
"""
from . import _capi  # noqa: F401

__all__ = ["_capi"]
