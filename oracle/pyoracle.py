"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, always as the checker of the HIP backend, never as a product path.
States are numpy complex128 arrays modified in place.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libqoa_oracle.so"

P, D, U32, U64, I32 = C.c_void_p, C.c_double, C.c_uint32, C.c_uint64, C.c_int
DP = C.POINTER(C.c_double)


class OrcRng(C.Structure):
    _fields_ = [("key", C.c_uint32 * 2), ("ctr", C.c_uint64), ("buf", C.c_uint32 * 4),
                ("have", C.c_int)]


_SIGS = {
    "orc_hadamard": [P, U64, D, D, U64], "orc_x": [P, U64, U64], "orc_y": [P, U64, U64],
    "orc_z": [P, U64, U64], "orc_phase_factor": [P, U64, U64, D, D], "orc_t": [P, U64, U64],
    "orc_s": [P, U64, U64], "orc_phaseshift": [P, U64, U64, D], "orc_reset": [P, U64, U64],
    "orc_swap": [P, U64, U64, U64], "orc_controlled_swap": [P, U64, U64, U64, U64],
    "orc_matrix2": [P, U64, U64, P], "orc_rx": [P, U64, U64, D], "orc_ry": [P, U64, U64, D],
    "orc_rz": [P, U64, U64, D], "orc_cnot": [P, U64, U64, U64], "orc_cz": [P, U64, U64, U64],
    "orc_cphase": [P, U64, U64, U64, D], "orc_reset_all": [P, U64],
    "orc_custom_unitary": [P, U64, U64, P], "orc_two_qubit_unitary": [P, U64, U64, U64, P],
    "orc_toffoli": [P, U64, U64, U64, U64], "orc_controlled_y": [P, U64, U64, U64],
    "orc_normalize": [P, U64], "orc_measure": [P, U64, U64, D, DP, DP],
    "orc_probabilities": [P, U64, P], "orc_pauli_codes": [P, U64, P, U32],
    "orc_run_gates": [P, U64, P, U32, P, U32],
    "orc_run_gates_ordered": [P, U64, P, P, U32, P, U32],
    "orc_philox4x32_10": [P, P, P], "orc_splitmix64_next": [C.POINTER(U64)],
    "orc_rng_seed": [C.POINTER(OrcRng), U64], "orc_rng_u64": [C.POINTER(OrcRng)],
    "orc_rng_f64": [C.POINTER(OrcRng)],
    "orc_qs_apply": [P, U64, U32, P, P, U32, I32, D, C.POINTER(OrcRng)],
    "orc_qs_measure": [P, U64, U32, U32, I32, D, C.POINTER(OrcRng), C.POINTER(I32)],
    "orc_qs_apply_noise": [P, U64, U32, I32, D, C.POINTER(OrcRng)],
    "orc_traj_pauli": [U64, U32, U32, U32, D],
    "orc_run_trajectory": [P, U64, P, U32, P, U32, U64, U32, D, D],
    "orc_run_trajectory_ordered": [P, U64, P, P, U32, P, U32, U64, U32, D, D],
    "orc_sample_uniform": [U64, U32, U32], "orc_sample": [P, U64, D],
    "orc_set_threads": [I32], "orc_run_gates_faithful": [P, U64, P, U32, P, U32, I32],
}
_RES = {"orc_normalize": I32, "orc_measure": I32, "orc_run_gates": I32,
        "orc_run_gates_ordered": I32, "orc_splitmix64_next": U64, "orc_rng_u64": U64,
        "orc_rng_f64": D, "orc_norm_sqr": D, "orc_qs_apply": I32, "orc_qs_measure": I32,
        "orc_traj_pauli": U32, "orc_run_trajectory": I32, "orc_run_trajectory_ordered": I32,
        "orc_sample_uniform": D, "orc_sample": U64, "orc_run_gates_faithful": I32}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `python oracle/build.py`")
        L = C.CDLL(str(LIB_PATH))
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RES.get(name, None)
        L.orc_norm_sqr.argtypes = [P, U64]
        L.orc_norm_sqr.restype = D
        _lib = L
    return _lib


def _p(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def state(n: int) -> np.ndarray:
    a = np.zeros(1 << n, dtype=np.complex128)
    a[0] = 1.0
    return a


class Oracle:
    """Thin call-through: o.hadamard(a, nf, q) etc. mutate `a` in place."""

    def __getattr__(self, name):
        f = getattr(lib(), "orc_" + name)

        def call(a, *args):
            if isinstance(a, np.ndarray):
                assert a.dtype == np.complex128 and a.flags["C_CONTIGUOUS"]
                conv = []
                for x in args:
                    if isinstance(x, np.ndarray):
                        conv.append(_p(np.ascontiguousarray(x)))
                    elif isinstance(x, complex):
                        conv.extend([x.real, x.imag])
                    else:
                        conv.append(x)
                return f(_p(a), a.size, *conv)
            return f(a, *args)

        return call


orc = Oracle()


def run_gates(a: np.ndarray, gates: np.ndarray, flags: int = 0, matrices=None, order=None) -> None:
    g = np.ascontiguousarray(gates)
    m = None if matrices is None else _p(np.ascontiguousarray(matrices, dtype=np.complex128))
    if order is None:
        rc = lib().orc_run_gates(_p(a), a.size, _p(g), len(g), m, flags)
    else:
        o = np.ascontiguousarray(order, dtype=np.uint32)
        rc = lib().orc_run_gates_ordered(_p(a), a.size, _p(g), _p(o), len(g), m, flags)
    if rc != 0:
        raise ValueError("oracle rejected a gate")


def set_threads(n: int) -> None:
    """Threads for the gate loops (the all-core CPU baseline); 1 restores the
    sequential restatement."""
    lib().orc_set_threads(int(n))


def run_gates_faithful(a: np.ndarray, gates: np.ndarray, threads: int, flags: int = 0,
                       matrices=None) -> None:
    """One-thread sweep plus a `threads`-thread normalize after every gate."""
    g = np.ascontiguousarray(gates)
    m = None if matrices is None else _p(np.ascontiguousarray(matrices, dtype=np.complex128))
    if lib().orc_run_gates_faithful(_p(a), a.size, _p(g), len(g), m, flags, int(threads)) != 0:
        raise ValueError("oracle rejected a gate")


def measure(a: np.ndarray, q: int, r: float) -> tuple[int, float, float]:
    p0, p1 = C.c_double(), C.c_double()
    o = lib().orc_measure(_p(a), a.size, q, r, C.byref(p0), C.byref(p1))
    return o, p0.value, p1.value


def norm_sqr(a: np.ndarray) -> float:
    return lib().orc_norm_sqr(_p(a), a.size)


def probabilities(a: np.ndarray) -> np.ndarray:
    out = np.empty(a.size, dtype=np.float64)
    lib().orc_probabilities(_p(a), a.size, _p(out))
    return out


def philox(ctr, key) -> list[int]:
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(o))
    return [int(x) for x in o]


def splitmix64(seed: int, count: int) -> list[int]:
    st = C.c_uint64(seed)
    return [int(lib().orc_splitmix64_next(C.byref(st))) for _ in range(count)]


class QS:
    """QuantumState semantics on the CPU (noise, normalize after every gate)."""

    def __init__(self, n: int, noise_kind: int = 0, p: float = 0.0, seed: int = 0,
                 flags: int = 0):
        self.n = max(n, 1)
        self.a = state(self.n)
        self.kind, self.p, self.flags = noise_kind, p, flags
        self.rng = OrcRng()
        lib().orc_rng_seed(C.byref(self.rng), seed)

    def apply(self, g: np.ndarray, matrices=None) -> int:
        """One gate (a GATE_DTYPE array of length 1) with QuantumState semantics."""
        g = np.ascontiguousarray(np.asarray(g).reshape(-1)[:1])
        m = None if matrices is None else _p(np.ascontiguousarray(matrices, dtype=np.complex128))
        return lib().orc_qs_apply(_p(self.a), self.a.size, self.n, _p(g), m, self.flags,
                                  self.kind, self.p, C.byref(self.rng))

    def measure(self, q: int) -> int:
        o = C.c_int()
        rc = lib().orc_qs_measure(_p(self.a), self.a.size, self.n, q, self.kind, self.p,
                                  C.byref(self.rng), C.byref(o))
        if rc != 0:
            raise ValueError(f"qubit index {q} out of bounds")
        return o.value


def run_trajectory(a: np.ndarray, gates: np.ndarray, seed: int, traj: int, p1: float, p2: float,
                   flags: int = 0, matrices=None, order=None) -> None:
    g = np.ascontiguousarray(gates)
    m = None if matrices is None else _p(np.ascontiguousarray(matrices, dtype=np.complex128))
    if order is None:
        rc = lib().orc_run_trajectory(_p(a), a.size, _p(g), len(g), m, flags, seed, traj, p1, p2)
    else:
        o = np.ascontiguousarray(order, dtype=np.uint32)
        rc = lib().orc_run_trajectory_ordered(_p(a), a.size, _p(g), _p(o), len(g), m, flags,
                                              seed, traj, p1, p2)
    if rc != 0:
        raise ValueError("oracle rejected a gate")


def traj_pauli(seed: int, traj: int, gate_idx: int, qubit: int, p: float) -> int:
    return int(lib().orc_traj_pauli(seed, traj, gate_idx, qubit, p))
