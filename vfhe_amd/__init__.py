"""vfhe_amd: an MI355X-native `hip` engine for vfhe's RNS polynomial hot path.

The engine itself is the C-ABI library libvfhe_hip.so (include/vfhe_hip.h),
built from csrc/.  This package holds its ctypes binding (`_native`), numpy
helpers (`runtime`), the host-side mirror of the reference's Python layers
that drive the path (`arith`, `mlwe`, `ckks`), and the engine picker
(`engine`).
"""
from . import _native  # noqa: F401
from .runtime import Pool  # noqa: F401
