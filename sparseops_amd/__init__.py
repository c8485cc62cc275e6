"""sparseops_amd: MI355X-native SpMV and LeSpGEMM (SparseOps hot path) behind a C ABI.

The compute path is libsparseops_hip.so (hand-written gfx950 HIP kernels), loaded through
ctypes by sparseops_amd._lib; importing this package raises if the library is missing.
"""
from . import _lib  # noqa: F401  (fails loudly when the HIP library is not built)

__all__ = ["_lib"]
