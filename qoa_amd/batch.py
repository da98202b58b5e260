"""Batched execution of many independent circuits (one launch per pass round)."""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _native as N
from ._native import check, load, ptr
from .state import gates_array


class Batch:
    """Jobs given as (n_qubits, gates) pairs; every job starts from |0...0>.

    precision N.QOA_C128: fused FP64 register passes (noise allowed);
    N.QOA_C64: the FP32-fused form, dense 5-qubit blocks on matrix cores.
    flags may hold N.QOA_BATCH_AUTOTUNE (tiles chosen per qubit count)."""

    def __init__(self, jobs: Sequence[tuple[int, np.ndarray]], tile_qubits: int = 12,
                 low_qubits: int = 0, flags: int = 0, trajectories: int = 1,
                 matrices: np.ndarray | None = None, p1: float = 0.0, p2: float = 0.0,
                 noise_seed: int | Sequence[int] = 0, precision: int = N.QOA_C128):
        self._lib = load()
        arrays = [gates_array(g) for _, g in jobs]
        self.gates = np.ascontiguousarray(np.concatenate(arrays)) if arrays else \
            np.zeros(0, dtype=N.GATE_DTYPE)
        descs = (N.QoaJobDesc * len(jobs))()
        off = 0
        self.n = []
        for i, ((n, _), g) in enumerate(zip(jobs, arrays)):
            seed = noise_seed if isinstance(noise_seed, int) else int(noise_seed[i])
            descs[i] = N.QoaJobDesc(int(n), len(g), off, seed, trajectories, precision, p1, p2)
            off += len(g)
            self.n.append(int(n))
        self.n_gates = [len(g) for g in arrays]
        m = None if matrices is None else np.ascontiguousarray(matrices, dtype=np.complex128)
        opts = N.QoaFuseOpts(tile_qubits, low_qubits, flags, 0)
        h = C.c_void_p()
        check(self._lib.qoa_batch_create(descs, len(jobs), ptr(self.gates), len(self.gates),
                                         None if m is None else ptr(m), C.byref(opts),
                                         C.byref(h)))
        self._h = h
        self.trajectories = trajectories

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.qoa_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self) -> None:
        check(self._lib.qoa_batch_run(self._h))

    def synchronize(self) -> float:
        ms = C.c_double()
        check(self._lib.qoa_batch_synchronize(self._h, C.byref(ms)))
        return ms.value

    def info(self) -> dict:
        r, a, b = C.c_uint32(), C.c_uint64(), C.c_double()
        check(self._lib.qoa_batch_info(self._h, C.byref(r), C.byref(a), C.byref(b)))
        return {"rounds": r.value, "total_amps": a.value, "alg_bytes": b.value}

    def download(self, job: int, traj: int = 0) -> np.ndarray:
        out = np.empty(1 << self.n[job], dtype=np.complex128)
        check(self._lib.qoa_batch_download(self._h, job, traj, ptr(out), out.size))
        return out

    def job_plan(self, job: int) -> tuple[int, int]:
        """(tile qubits, passes) of a C128 job; (5, blocks) of a C64 job."""
        t, n = C.c_uint32(), C.c_uint32()
        f = self._lib.qoa_batch_job_plan
        f.argtypes, f.restype = N._SIGS["qoa_batch_job_plan"]
        check(f(self._h, job, C.byref(t), C.byref(n)))
        return t.value, n.value

    def order(self, job: int) -> np.ndarray:
        o = np.empty(self.n_gates[job], dtype=np.uint32)
        check(self._lib.qoa_batch_order(self._h, job, ptr(o), o.size))
        return o
