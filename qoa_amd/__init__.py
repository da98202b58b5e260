"""qoa_amd — MI355X (gfx950) state-vector backend for QOA's gate-application path.

The compute path is libqoa_hip.so (hand-written HIP kernels behind the C ABI of
include/qoa_hip.h); this package is the host-side mirror of the reference's
state API plus the workload generators used by tests and benchmarks.
"""
from ._native import GATE_DTYPE, OP, QoaError, load  # noqa: F401
from .state import (NoiseConfig, QuantumState, StateVector, gates_array, plan_blocks,  # noqa: F401
                    plan_fused)

__all__ = ["GATE_DTYPE", "OP", "QoaError", "load", "NoiseConfig", "QuantumState", "StateVector",
           "gates_array", "plan_blocks", "plan_fused"]
