"""atomiverse_amd — MI355X (gfx950) classical-MD force and integration backend.

The backend itself is libatv_hip.so (hand-written HIP kernels behind the C ABI
in include/atvh.h) and the hip:: C++ components in include/atomiverse/hip/.
This Python package only binds the C ABI for tests and benchmarks; it has no
CPU fallback and raises if the HIP library is not built.
"""
from . import _abi  # noqa: F401
from ._abi import PHASE_FINAL, PHASE_INITIAL, PREC_FP64, PREC_MIXED, HipError, load  # noqa: F401
from .system import Batch, Context, Server, System, segment_bytes  # noqa: F401

__all__ = ["Batch", "Context", "Server", "System", "segment_bytes", "HipError", "load", "PREC_FP64", "PREC_MIXED", "PHASE_INITIAL", "PHASE_FINAL"]
