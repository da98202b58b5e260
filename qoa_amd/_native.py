"""ctypes binding of libqoa_hip.so (the C ABI declared in include/qoa_hip.h).

The library is built in-tree by ``python -m qoa_amd.build``. Loading fails
loudly when it is missing: there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libqoa_hip.so"
HEADER = PKG.parent / "include" / "qoa_hip.h"

QOA_OK = 0
QOA_ERR_HANDLE = -1
QOA_ERR_QUBIT = -2
QOA_ERR_OOM = -3
QOA_ERR_HIP = -4
QOA_ERR_ARG = -5
QOA_ERR_ADMISSION = -6
QOA_ERR_UNSUPPORTED = -7

QOA_C128 = 0
QOA_C64 = 1
QOA_CZ_REF_MASKS = 1
QOA_FUSE_LDS_ONLY = 2
QOA_FUSE_FMA = 4
QOA_FUSE_MERGE_DIAG = 16
QOA_FUSE_SPECIALIZE = 64
QOA_FUSE_FACTOR = 128
QOA_FUSE_RELABEL = 256
QOA_BATCH_AUTOTUNE = 32
QOA_BLOCK_NO_RENORM = 8

NOISE_NONE, NOISE_IDEAL, NOISE_RANDOM, NOISE_FIXED = 0, 1, 2, 3

# op codes (reference bytecode bytes where they exist, src/instructions.rs:631-830)
OP = dict(H=0x05, Z=0x06, X=0x07, PHASESHIFT=0x08, RESET=0x0A, SWAP=0x0B, CSWAP=0x0C, T=0x0D,
          S=0x0E, RZ=0x0F, CNOT=0x17, RESET_ALL=0x1C, CZ=0x1E, RX=0x22, RY=0x23, PHASE=0x24,
          MEASURE=0x32, CPHASE=0x35, Y=0xE1, U2=0xE2, U4=0xE3, TOFFOLI=0xE4, CY=0xE5)

GATE_DTYPE = np.dtype([("op", "<u4"), ("q0", "<u4"), ("q1", "<u4"), ("q2", "<u4"),
                       ("theta", "<f8"), ("aux", "<u8")], align=True)
assert GATE_DTYPE.itemsize == 32


class QoaGate(C.Structure):
    _fields_ = [("op", C.c_uint32), ("q0", C.c_uint32), ("q1", C.c_uint32), ("q2", C.c_uint32),
                ("theta", C.c_double), ("aux", C.c_uint64)]


class QoaFuseOpts(C.Structure):
    _fields_ = [("tile_qubits", C.c_uint32), ("low_qubits", C.c_uint32), ("flags", C.c_uint32),
                ("reserved", C.c_uint32)]


class QoaBlockOpts(C.Structure):
    _fields_ = [("block_qubits", C.c_uint32), ("flags", C.c_uint32)]


class QoaJobDesc(C.Structure):
    _fields_ = [("n_qubits", C.c_uint32), ("n_gates", C.c_uint32), ("gate_offset", C.c_uint64),
                ("noise_seed", C.c_uint64), ("trajectories", C.c_uint32),
                ("precision", C.c_uint32), ("p1", C.c_double), ("p2", C.c_double)]


class QoaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


P = C.c_void_p
D = C.c_double
U32 = C.c_uint32
U64 = C.c_uint64
I32 = C.c_int
DP = C.POINTER(C.c_double)

_SIGS = {
    "qoa_version": ([], I32),
    "qoa_last_error": ([], C.c_char_p),
    "qoa_device_info": ([C.c_char_p, C.c_size_t, C.POINTER(U64), C.POINTER(I32)], I32),
    "qoa_set_option": ([C.c_char_p, C.c_int64], I32),
    "qoa_state_create": ([U32, U32, C.POINTER(P)], I32),
    "qoa_state_destroy": ([P], I32),
    "qoa_state_num_qubits": ([P, C.POINTER(U32)], I32),
    "qoa_state_upload": ([P, P, U64], I32),
    "qoa_state_download": ([P, P, U64], I32),
    "qoa_state_synchronize": ([P], I32),
    "qoa_state_copy": ([P, P], I32),
    "qoa_apply_hadamard": ([P, D, D, U32], I32),
    "qoa_apply_x": ([P, U32], I32),
    "qoa_apply_y": ([P, U32], I32),
    "qoa_apply_z": ([P, U32], I32),
    "qoa_apply_t": ([P, U32], I32),
    "qoa_apply_s": ([P, U32], I32),
    "qoa_apply_phaseshift": ([P, U32, D], I32),
    "qoa_apply_reset": ([P, U32], I32),
    "qoa_apply_swap": ([P, U32, U32], I32),
    "qoa_apply_controlled_swap": ([P, U32, U32, U32], I32),
    "qoa_apply_rx": ([P, U32, D], I32),
    "qoa_apply_ry": ([P, U32, D], I32),
    "qoa_apply_rz": ([P, U32, D], I32),
    "qoa_apply_cnot": ([P, U32, U32], I32),
    "qoa_apply_cz": ([P, U64, U64], I32),
    "qoa_apply_controlled_phase_rotation": ([P, U64, U64, D], I32),
    "qoa_apply_reset_all": ([P], I32),
    "qoa_apply_matrix2": ([P, U32, P], I32),
    "qoa_apply_matrix4": ([P, U32, U32, P], I32),
    "qoa_apply_toffoli": ([P, U32, U32, U32], I32),
    "qoa_apply_controlled_y": ([P, U32, U32], I32),
    "qoa_norm_sqr": ([P, DP], I32),
    "qoa_normalize": ([P, DP, C.POINTER(I32)], I32),
    "qoa_measure": ([P, U32, D, C.POINTER(I32), DP, DP], I32),
    "qoa_probabilities": ([P, P, U64], I32),
    "qoa_apply_pauli_codes": ([P, P, U32], I32),
    "qoa_run_gates": ([P, P, U32, P, U32], I32),
    "qoa_validate": ([P, C.POINTER(I32), C.POINTER(I32), C.POINTER(D)], I32),
    "qoa_run_trajectory": ([P, P, U32, P, U32, U64, U32, D, D, P], I32),
    "qoa_run_fused": ([P, P, U32, P, P], I32),
    "qoa_plan_fused": ([U32, P, U32, P, P, C.POINTER(U32)], I32),
    "qoa_tile_op_size": ([], I32),
    "qoa_run_blocks": ([P, P, U32, P, C.POINTER(QoaBlockOpts), C.POINTER(U32)], I32),
    "qoa_plan_blocks": ([U32, P, U32, P, C.POINTER(QoaBlockOpts), P, P, U32, C.POINTER(U32)],
                        I32),
    "qoa_state_compare": ([P, P, DP, DP], I32),
    "qoa_plan_pool": ([U32, P, U32, P, P, P, U32, C.POINTER(U32)], I32),
    "qoa_plan_store_positions": ([U32, P, U32, P, P, P, U32, C.POINTER(U32)], I32),
    "qoa_specialize_compile": ([U32, P, U32, P, P, C.c_char_p], I32),
    "qoa_specialize_info": ([C.POINTER(U32), C.POINTER(U32), DP], I32),
    "qoa_plan_program": ([U32, P, U32, P, P, P, U32, P, P, U32, C.POINTER(U32), C.POINTER(U32),
                          C.POINTER(U32)], I32),
    "qoa_timing_enable": ([P, I32], I32),
    "qoa_timing_read": ([P, C.c_char_p, DP, C.POINTER(U64), DP], I32),
    "qoa_qs_set_noise": ([P, I32, D], I32),
    "qoa_qs_seed_rng": ([P, U64], I32),
    "qoa_qs_set_flags": ([P, U32], I32),
    "qoa_qs_apply": ([P, P, P], I32),
    "qoa_qs_measure": ([P, U32, C.POINTER(I32)], I32),
    "qoa_qs_measure_all": ([P, P, U32], I32),
    "qoa_qs_apply_noise": ([P], I32),
}

_lib = None


def header_functions() -> list[str]:
    """Names of every function declared in include/qoa_hip.h."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(qoa_\w+)\s*\(", text, flags=re.M)))


def load(path: Path | None = None) -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python -m qoa_amd.build` "
                           "(the HIP backend has no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (args, res) in _SIGS.items():
        f = getattr(lib, name, None)
        if f is None:
            continue
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def check(rc: int) -> int:
    if rc < 0:
        msg = load().qoa_last_error().decode(errors="replace")
        raise QoaError(rc, msg)
    return rc


def ptr(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


_SIGS.update({
    "qoa_batch_create": ([C.POINTER(QoaJobDesc), U32, P, U64, P, P, C.POINTER(P)], I32),
    "qoa_batch_run": ([P], I32),
    "qoa_batch_synchronize": ([P, DP], I32),
    "qoa_batch_info": ([P, C.POINTER(U32), C.POINTER(U64), DP], I32),
    "qoa_batch_download": ([P, U32, U32, P, U64], I32),
    "qoa_batch_order": ([P, U32, P, U32], I32),
    "qoa_batch_job_plan": ([P, U32, C.POINTER(U32), C.POINTER(U32)], I32),
    "qoa_batch_destroy": ([P], I32),
})


class QoaServerOpts(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("tile_qubits", C.c_uint32),
                ("planner_threads", C.c_uint32), ("max_items", C.c_uint32),
                ("max_ops", C.c_uint32), ("max_jobs", C.c_uint32), ("flags", C.c_uint32)]


class QoaJobResult(C.Structure):
    _fields_ = [("job_id", C.c_uint64), ("status", C.c_int32), ("n_qubits", C.c_uint32),
                ("trajectories", C.c_uint32), ("n_passes", C.c_uint32),
                ("n_samples", C.c_uint32), ("t_arrival", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double)]


class QoaServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("steps", C.c_uint64), ("arena_bytes", C.c_uint64), ("arena_peak", C.c_uint64),
                ("fatal", C.c_int32)]


P, U32, U64, D, I32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_double, C.c_int
_SIGS.update({
    "qoa_server_create": ([C.POINTER(QoaServerOpts), C.POINTER(P)], I32),
    "qoa_server_submit": ([P, C.POINTER(QoaJobDesc), P, P, U32, D, C.POINTER(U64)], I32),
    "qoa_server_now": ([P], D),
    "qoa_server_wait": ([P, U64, D, C.POINTER(QoaJobResult), P, U32], I32),
    "qoa_server_drain": ([P, D], I32),
    "qoa_server_get_stats": ([P, C.POINTER(QoaServerStats)], I32),
    "qoa_server_destroy": ([P], I32),
    "qoa_sample": ([P, U64, U32, P], I32),
    "qoa_sample_noalloc": ([P, U64, U32, P], I32),
})


