"""Python handle on the continuous-batching serving engine (server.hip)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from ._native import check, load, ptr
from .state import gates_array


from ._native import QoaJobResult, QoaServerOpts, QoaServerStats  # noqa: E402


@dataclass
class JobResult:
    job_id: int
    status: int
    n_qubits: int
    trajectories: int
    t_arrival: float
    t_admit: float
    t_done: float
    samples: np.ndarray

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


class Server:
    def __init__(self, budget_bytes: int = 0, tile_qubits: int = 12, planner_threads: int = 4,
                 max_jobs: int = 4096, flags: int = 0):
        self._lib = load()
        for name, (args, res) in N._SIGS.items():
            if name.startswith("qoa_server") or name == "qoa_sample":
                f = getattr(self._lib, name)
                f.argtypes, f.restype = args, res
        o = QoaServerOpts(budget_bytes, tile_qubits, planner_threads, 0, 0, max_jobs, flags)
        h = C.c_void_p()
        check(self._lib.qoa_server_create(C.byref(o), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.qoa_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self) -> float:
        return self._lib.qoa_server_now(self._h)

    def submit(self, n_qubits: int, gates, trajectories: int = 1, p1: float = 0.0,
               p2: float = 0.0, noise_seed: int = 0, arrival: float | None = None) -> int:
        g = gates_array(gates)
        d = N.QoaJobDesc(n_qubits, len(g), 0, noise_seed, trajectories, N.QOA_C128, p1, p2)
        jid = C.c_uint64()
        t = self.now() if arrival is None else arrival
        rc = self._lib.qoa_server_submit(self._h, C.byref(d), ptr(g), None, 0, t, C.byref(jid))
        if rc != N.QOA_ERR_ADMISSION:
            check(rc)
        return jid.value

    def wait(self, job_id: int, timeout: float = 60.0) -> JobResult:
        r = QoaJobResult()
        buf = np.zeros(4096, dtype=np.uint64)
        rc = self._lib.qoa_server_wait(self._h, job_id, timeout, C.byref(r), ptr(buf), buf.size)
        if rc not in (N.QOA_OK, N.QOA_ERR_ADMISSION) and r.job_id != job_id:
            check(rc)
        return JobResult(r.job_id, r.status, r.n_qubits, r.trajectories, r.t_arrival, r.t_admit,
                         r.t_done, buf[: r.n_samples].copy())

    def drain(self, timeout: float = 600.0) -> None:
        check(self._lib.qoa_server_drain(self._h, timeout))

    def stats(self) -> dict:
        s = QoaServerStats()
        check(self._lib.qoa_server_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in QoaServerStats._fields_}


def sample_state(sv, seed: int, count: int) -> np.ndarray:
    """Z-basis samples of a StateVector (fixed-order device CDF)."""
    lib = load()
    f = lib.qoa_sample
    f.argtypes, f.restype = N._SIGS["qoa_sample"]
    out = np.zeros(count, dtype=np.uint64)
    check(f(sv.handle, seed, count, ptr(out)))
    return out
