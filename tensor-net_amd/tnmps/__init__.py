"""MI355X-native Schmidt-form (Gamma-Lambda) MPS update path.

``MPS`` mirrors tensor-net's ``MPS<Q, C64>`` operator API
(tensor-net/lib/mps.rs) over the C ABI of ``libtnmps.so``; all arithmetic
runs in hand-written HIP kernels for gfx950.
"""
from .mps import MPS, Engine, MPSError, MPSWarning  # noqa: F401
from . import workload  # noqa: F401

__all__ = ["MPS", "Engine", "MPSError", "MPSWarning", "workload"]
