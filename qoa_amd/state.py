"""Host-side mirror of QOA's state API over the HIP C ABI.

``StateVector`` exposes the layer-1 sweeps of src/vectorization.rs on a
device-resident state; ``QuantumState`` mirrors the wrapper methods of
src/runtime/quantum_state.rs (bounds checks, noise, renormalisation after
every gate, measurement) with the same names and argument meaning.
"""
from __future__ import annotations

import ctypes as C
import math
import sys
from typing import Iterable, Optional

import numpy as np

from . import _native as N
from ._native import OP, QoaError, check, load, ptr


def gates_array(gates: Iterable) -> np.ndarray:
    """Turns (op, q0, q1, q2, theta[, aux]) tuples or a GATE_DTYPE array into a
    contiguous GATE_DTYPE array."""
    if isinstance(gates, np.ndarray) and gates.dtype == N.GATE_DTYPE:
        return np.ascontiguousarray(gates)
    rows = []
    for g in gates:
        g = tuple(g)
        op = OP[g[0]] if isinstance(g[0], str) else int(g[0])
        q0 = int(g[1]) if len(g) > 1 else 0
        q1 = int(g[2]) if len(g) > 2 else 0
        q2 = int(g[3]) if len(g) > 3 else 0
        th = float(g[4]) if len(g) > 4 else 0.0
        aux = int(g[5]) if len(g) > 5 else 0
        rows.append((op, q0, q1, q2, th, aux))
    return np.array(rows, dtype=N.GATE_DTYPE)


class StateVector:
    """A 2^n-amplitude state held in HBM (complex128 by default)."""

    def __init__(self, n_qubits: int, precision: int = N.QOA_C128):
        self._lib = load()
        h = C.c_void_p()
        check(self._lib.qoa_state_create(int(n_qubits), int(precision), C.byref(h)))
        self._h = h
        n = C.c_uint32()
        check(self._lib.qoa_state_num_qubits(h, C.byref(n)))
        self.n = int(n.value)
        self.precision = precision

    # -- lifecycle -------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.qoa_state_destroy(self._h)
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

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def length(self) -> int:
        return 1 << self.n

    def upload(self, amps: np.ndarray) -> None:
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        if a.size != self.length:
            raise ValueError(f"expected {self.length} amplitudes, got {a.size}")
        check(self._lib.qoa_state_upload(self._h, ptr(a), a.size))

    def download(self) -> np.ndarray:
        out = np.empty(self.length, dtype=np.complex128)
        check(self._lib.qoa_state_download(self._h, ptr(out), out.size))
        return out

    def synchronize(self) -> None:
        check(self._lib.qoa_state_synchronize(self._h))

    def copy_from(self, other: "StateVector") -> None:
        check(self._lib.qoa_state_copy(self._h, other._h))

    # -- layer 1 (src/vectorization.rs) ----------------------------------------
    def apply_hadamard(self, mask_bit: int, norm_factor: complex = 1.0 / math.sqrt(2.0)) -> None:
        nf = complex(norm_factor)
        check(self._lib.qoa_apply_hadamard(self._h, nf.real, nf.imag, mask_bit))

    def apply_x(self, q: int) -> None: check(self._lib.qoa_apply_x(self._h, q))
    def apply_y(self, q: int) -> None: check(self._lib.qoa_apply_y(self._h, q))
    def apply_z(self, q: int) -> None: check(self._lib.qoa_apply_z(self._h, q))
    def apply_t(self, q: int) -> None: check(self._lib.qoa_apply_t(self._h, q))
    def apply_s(self, q: int) -> None: check(self._lib.qoa_apply_s(self._h, q))

    def apply_phaseshift(self, q: int, angle: float) -> None:
        check(self._lib.qoa_apply_phaseshift(self._h, q, angle))

    def apply_reset(self, q: int) -> None: check(self._lib.qoa_apply_reset(self._h, q))
    def apply_reset_all(self) -> None: check(self._lib.qoa_apply_reset_all(self._h))

    def apply_swap(self, q1: int, q2: int) -> None:
        check(self._lib.qoa_apply_swap(self._h, q1, q2))

    def apply_controlled_swap(self, c: int, t1: int, t2: int) -> None:
        check(self._lib.qoa_apply_controlled_swap(self._h, c, t1, t2))

    def apply_rx(self, q: int, angle: float) -> None: check(self._lib.qoa_apply_rx(self._h, q, angle))
    def apply_ry(self, q: int, angle: float) -> None: check(self._lib.qoa_apply_ry(self._h, q, angle))
    def apply_rz(self, q: int, angle: float) -> None: check(self._lib.qoa_apply_rz(self._h, q, angle))

    def apply_cnot(self, control: int, target: int) -> None:
        check(self._lib.qoa_apply_cnot(self._h, control, target))

    def apply_cz(self, control_mask: int, target_mask: int) -> None:
        check(self._lib.qoa_apply_cz(self._h, control_mask, target_mask))

    def apply_controlled_phase_rotation(self, control_mask: int, target_mask: int,
                                        angle: float) -> None:
        check(self._lib.qoa_apply_controlled_phase_rotation(self._h, control_mask, target_mask,
                                                            angle))

    def apply_matrix2(self, q: int, m: np.ndarray) -> None:
        mm = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(4))
        check(self._lib.qoa_apply_matrix2(self._h, q, ptr(mm)))

    def apply_matrix4(self, q1: int, q2: int, m: np.ndarray) -> None:
        mm = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(16))
        check(self._lib.qoa_apply_matrix4(self._h, q1, q2, ptr(mm)))

    def apply_toffoli(self, c1: int, c2: int, t: int) -> None:
        check(self._lib.qoa_apply_toffoli(self._h, c1, c2, t))

    def apply_controlled_y(self, c: int, t: int) -> None:
        check(self._lib.qoa_apply_controlled_y(self._h, c, t))

    def apply_pauli_codes(self, codes: Iterable[int]) -> None:
        c = np.ascontiguousarray(np.asarray(list(codes), dtype=np.uint8))
        check(self._lib.qoa_apply_pauli_codes(self._h, ptr(c), c.size))

    # -- reductions ------------------------------------------------------------
    def norm_sqr(self) -> float:
        v = C.c_double()
        check(self._lib.qoa_norm_sqr(self._h, C.byref(v)))
        return v.value

    def normalize(self) -> tuple[float, bool]:
        v = C.c_double()
        r = C.c_int()
        check(self._lib.qoa_normalize(self._h, C.byref(v), C.byref(r)))
        return v.value, bool(r.value)

    def measure_with(self, q: int, r: float) -> tuple[int, float, float]:
        o, p0, p1 = C.c_int(), C.c_double(), C.c_double()
        check(self._lib.qoa_measure(self._h, q, r, C.byref(o), C.byref(p0), C.byref(p1)))
        return o.value, p0.value, p1.value

    def probabilities(self) -> np.ndarray:
        out = np.empty(self.length, dtype=np.float64)
        check(self._lib.qoa_probabilities(self._h, ptr(out), out.size))
        return out

    # -- circuits ----------------------------------------------------------------
    def run_gates(self, gates, flags: int = 0, matrices: Optional[np.ndarray] = None) -> None:
        g = gates_array(gates)
        m = None if matrices is None else np.ascontiguousarray(matrices, dtype=np.complex128)
        check(self._lib.qoa_run_gates(self._h, ptr(g), len(g), None if m is None else ptr(m),
                                      flags))

    def run_trajectory(self, gates, seed: int, traj: int, p1: float, p2: float, flags: int = 0,
                       matrices: Optional[np.ndarray] = None, order=None) -> None:
        """One noisy trajectory as a plain per-gate loop (each gate's sweep,
        then the drawn Pauli sweeps on its operands)."""
        g = gates_array(gates)
        m = None if matrices is None else np.ascontiguousarray(matrices, dtype=np.complex128)
        o = None if order is None else np.ascontiguousarray(order, dtype=np.uint32)
        check(self._lib.qoa_run_trajectory(self._h, ptr(g), len(g), None if m is None else ptr(m),
                                           flags, seed, traj, p1, p2,
                                           None if o is None else ptr(o)))

    def run_fused(self, gates, tile_qubits: int = 0, low_qubits: int = 0, flags: int = 0,
                  matrices: Optional[np.ndarray] = None) -> None:
        g = gates_array(gates)
        m = None if matrices is None else np.ascontiguousarray(matrices, dtype=np.complex128)
        o = N.QoaFuseOpts(tile_qubits, low_qubits, flags, 0)
        check(self._lib.qoa_run_fused(self._h, ptr(g), len(g), None if m is None else ptr(m),
                                      C.byref(o)))

    def run_blocks(self, gates, block_qubits: int = 0, flags: int = 0,
                   matrices: Optional[np.ndarray] = None) -> int:
        """Dense 5-qubit blocks on FP32 matrix cores (complex64 states only);
        returns the number of blocks."""
        g = gates_array(gates)
        m = None if matrices is None else np.ascontiguousarray(matrices, dtype=np.complex128)
        o = N.QoaBlockOpts(block_qubits, flags)
        nb = C.c_uint32()
        check(self._lib.qoa_run_blocks(self._h, ptr(g), len(g), None if m is None else ptr(m),
                                       C.byref(o), C.byref(nb)))
        return int(nb.value)

    def compare(self, other: "StateVector") -> tuple[float, float]:
        """(max |a_i - b_i|, fidelity) computed on the device."""
        d, f = C.c_double(), C.c_double()
        check(self._lib.qoa_state_compare(self._h, other._h, C.byref(d), C.byref(f)))
        return d.value, f.value

    # -- timing ------------------------------------------------------------------
    def timing_enable(self, on: bool = True) -> None:
        check(self._lib.qoa_timing_enable(self._h, 1 if on else 0))

    def timing_read(self, kind: str) -> tuple[float, int, float]:
        ms, n, b = C.c_double(), C.c_uint64(), C.c_double()
        check(self._lib.qoa_timing_read(self._h, kind.encode(), C.byref(ms), C.byref(n),
                                        C.byref(b)))
        return ms.value, int(n.value), b.value


def plan_fused(n_qubits: int, gates, tile_qubits: int = 0, low_qubits: int = 0,
               flags: int = 0) -> tuple[np.ndarray, int]:
    """The fused planner's execution order of gate indices and its pass count."""
    lib = load()
    g = gates_array(gates)
    order = np.empty(len(g), dtype=np.uint32)
    np_ = C.c_uint32()
    o = N.QoaFuseOpts(tile_qubits, low_qubits, flags, 0)
    check(lib.qoa_plan_fused(n_qubits, ptr(g), len(g), C.byref(o), ptr(order), C.byref(np_)))
    return order, int(np_.value)


def plan_blocks(n_qubits: int, gates, block_qubits: int = 0, flags: int = 0,
                matrices: Optional[np.ndarray] = None) -> tuple[np.ndarray, np.ndarray]:
    """The dense-block fuser's output: qubits (blocks x 5, ascending) and the
    FP64 32x32 unitaries (blocks x 32 x 32), in execution order."""
    lib = load()
    g = gates_array(gates)
    m = None if matrices is None else np.ascontiguousarray(matrices, dtype=np.complex128)
    o = N.QoaBlockOpts(block_qubits, flags)
    nb = C.c_uint32()
    mp = None if m is None else ptr(m)
    check(lib.qoa_plan_blocks(n_qubits, ptr(g), len(g), mp, C.byref(o), None, None, 0,
                              C.byref(nb)))
    qs = np.zeros((nb.value, 5), dtype=np.uint32)
    us = np.zeros((nb.value, 32, 32), dtype=np.complex128)
    check(lib.qoa_plan_blocks(n_qubits, ptr(g), len(g), mp, C.byref(o), ptr(qs), ptr(us),
                              nb.value, C.byref(nb)))
    return qs, us


class NoiseConfig:
    """src/runtime/quantum_state.rs:15-24."""

    def __init__(self, kind: str, p: float = 0.0):
        assert kind in ("Random", "Fixed", "Ideal")
        self.kind, self.p = kind, float(p)

    @staticmethod
    def Random() -> "NoiseConfig": return NoiseConfig("Random")

    @staticmethod
    def Fixed(p: float) -> "NoiseConfig": return NoiseConfig("Fixed", p)

    @staticmethod
    def Ideal() -> "NoiseConfig": return NoiseConfig("Ideal")

    def code(self) -> int:
        return {"Ideal": N.NOISE_IDEAL, "Random": N.NOISE_RANDOM, "Fixed": N.NOISE_FIXED}[self.kind]


class QuantumState(StateVector):
    """src/runtime/quantum_state.rs:37-1212 over a device-resident state.

    Gate methods print the reference's error text and leave the state untouched
    when a qubit is out of range (quantum_state.rs:189-195); ``measure`` raises
    ``ValueError`` like the reference's ``Err`` (:400-402).
    """

    def __init__(self, n_qubits: int, noise_config: Optional[NoiseConfig] = None,
                 cz_ref_masks: bool = False):
        super().__init__(n_qubits)
        self.noise_config = None
        self.set_noise_config(noise_config)
        if cz_ref_masks:
            check(self._lib.qoa_qs_set_flags(self._h, N.QOA_CZ_REF_MASKS))

    def set_noise_config(self, cfg: Optional[NoiseConfig]) -> None:
        self.noise_config = cfg
        kind = N.NOISE_NONE if cfg is None else cfg.code()
        check(self._lib.qoa_qs_set_noise(self._h, kind, 0.0 if cfg is None else cfg.p))

    def get_noise_config(self) -> Optional[NoiseConfig]:
        return self.noise_config

    def seed_rng(self, seed: int) -> None:
        check(self._lib.qoa_qs_seed_rng(self._h, seed & 0xFFFFFFFFFFFFFFFF))

    def get_num_qubits(self) -> int:
        return self.n

    @property
    def amps(self) -> np.ndarray:
        return self.download()

    def _apply(self, op: str, q0=0, q1=0, q2=0, theta=0.0, matrix=None) -> None:
        g = gates_array([(op, q0, q1, q2, theta, 0)])
        m = None if matrix is None else np.ascontiguousarray(matrix, dtype=np.complex128)
        rc = self._lib.qoa_qs_apply(self._h, ptr(g), None if m is None else ptr(m))
        if rc == N.QOA_ERR_QUBIT:
            print(self._lib.qoa_last_error().decode(), file=sys.stderr)
            return
        check(rc)

    def reset_qubit(self, q): self._apply("RESET", q)
    def apply_reset_all(self): self._apply("RESET_ALL")
    def apply_h(self, q): self._apply("H", q)
    def apply_x(self, q): self._apply("X", q)
    def apply_y(self, q): self._apply("Y", q)
    def apply_phase_flip(self, q): self._apply("Z", q)
    def apply_t_gate(self, q): self._apply("T", q)
    def apply_s_gate(self, q): self._apply("S", q)
    def apply_phase_shift(self, q, angle): self._apply("PHASESHIFT", q, theta=angle)
    def apply_rx(self, q, angle): self._apply("RX", q, theta=angle)
    def apply_ry(self, q, angle): self._apply("RY", q, theta=angle)
    def apply_rz(self, q, angle): self._apply("RZ", q, theta=angle)
    def apply_cnot(self, c, t): self._apply("CNOT", c, t)
    def apply_cz(self, c, t): self._apply("CZ", c, t)
    def apply_controlled_phase(self, c, t, angle): self._apply("CPHASE", c, t, theta=angle)
    def apply_swap(self, q1, q2): self._apply("SWAP", q1, q2)
    def apply_controlled_swap(self, c, t1, t2): self._apply("CSWAP", c, t1, t2)
    def apply_toffoli(self, c1, c2, t): self._apply("TOFFOLI", c1, c2, t)
    def apply_controlled_y(self, c, t): self._apply("CY", c, t)

    def apply_custom_unitary(self, q: int, u) -> None:
        u = np.asarray(u, dtype=np.complex128).reshape(-1)
        if q >= self.n:
            raise ValueError(f"qubit index {q} out of bounds for {self.n}-qubit state "
                             "(apply_custom_unitary)")
        if u.size != 4:
            raise ValueError("unitary matrix must be 2x2 (4 elements)")
        self._apply("U2", q, matrix=u)

    def apply_two_qubit_unitary(self, q1: int, q2: int, u) -> None:
        u = np.asarray(u, dtype=np.complex128).reshape(-1)
        if q1 >= self.n or q2 >= self.n:
            raise ValueError(f"qubit indices ({q1}, {q2}) out of bounds for {self.n}-qubit state "
                             "(apply_two_qubit_unitary)")
        if q1 == q2:
            raise ValueError("qubits cannot be the same for two-qubit unitary.")
        if u.size != 16:
            raise ValueError("unitary matrix must be 4x4 (16 elements)")
        self._apply("U4", q1, q2, matrix=u)

    def measure(self, qubit_idx: int) -> int:
        if qubit_idx >= self.n:
            raise ValueError(f"qubit index {qubit_idx} out of bounds")
        o = C.c_int()
        check(self._lib.qoa_qs_measure(self._h, qubit_idx, C.byref(o)))
        return o.value

    def measure_all(self) -> list[int]:
        out = np.zeros(self.n, dtype=np.uint8)
        check(self._lib.qoa_qs_measure_all(self._h, ptr(out), self.n))
        return [int(x) for x in out]

    def get_probabilities(self) -> np.ndarray:
        return self.probabilities()

    def validate_state(self) -> None:
        """quantum_state.rs:135-167 (raises ValueError); one device pass, no
        copy of the state to the host."""
        nan, inf, s = C.c_int(), C.c_int(), C.c_double()
        check(self._lib.qoa_validate(self._h, C.byref(nan), C.byref(inf), C.byref(s)))
        if nan.value:
            raise ValueError("quantum state contains nan values.")
        if inf.value:
            raise ValueError("quantum state contains infinite values.")
        if abs(s.value - 1.0) > 1e-9:
            raise ValueError(f"quantum state is not normalized. norm squared: {s.value}")

    def get_state_vector(self) -> np.ndarray:
        return self.download()

    def set_state_vector(self, amps) -> None:
        a = np.asarray(amps, dtype=np.complex128)
        if a.size != self.length:
            raise ValueError(f"new amplitude vector length ({a.size}) does not match expected "
                             f"length ({self.length}) for {self.n} qubits")
        self.upload(a)
        self.normalize()
