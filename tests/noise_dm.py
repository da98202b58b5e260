"""Exact noisy outcome distributions by density matrices (numpy, test helper).

Independent of the oracle: every gate is its textbook matrix, every noise
event is the channel's definition written as a superoperator. The trajectory
channel of the serving workload (after a one-qubit gate each operand takes
X, Y or Z with total probability p1; after a gate on two or three qubits each
operand independently with p2; the Pauli is uniform over the three) maps one
qubit's density matrix as

    rho -> (1 - p) rho + p/3 (X rho X + Y rho Y + Z rho Z).

`mistake` selects a deliberately wrong channel, so a test can show that the
tolerance it uses would detect that mistake.
"""
from __future__ import annotations

import math

import numpy as np

from qoa_amd._native import OP

I2 = np.eye(2, dtype=np.complex128)
X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2)


def one_qubit_matrix(op: int, theta: float) -> np.ndarray:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    if op == OP["H"]:
        return H
    if op == OP["X"]:
        return X
    if op == OP["Y"]:
        return Y
    if op == OP["Z"]:
        return Z
    if op == OP["S"]:
        return np.diag([1, 1j]).astype(np.complex128)
    if op == OP["T"]:
        return np.diag([1, np.exp(1j * math.pi / 4)])
    if op in (OP["PHASE"], OP["PHASESHIFT"]):
        return np.diag([1, np.exp(1j * theta)])
    if op == OP["RX"]:
        return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)
    if op == OP["RY"]:
        return np.array([[c, -s], [s, c]], dtype=np.complex128)
    if op == OP["RZ"]:
        return np.diag([np.exp(-1j * theta / 2), np.exp(1j * theta / 2)])
    raise ValueError(f"no one-qubit matrix for op {op:#x}")


def full_unitary(n: int, g) -> np.ndarray:
    """The 2^n x 2^n matrix of one gate (qubit q = bit q of the index;
    CZ/CPHASE with textbook masks)."""
    op, q0, q1, q2, th = int(g["op"]), int(g["q0"]), int(g["q1"]), int(g["q2"]), float(g["theta"])
    N = 1 << n
    idx = np.arange(N)
    if op in (OP["CNOT"], OP["TOFFOLI"], OP["CSWAP"], OP["SWAP"]):
        perm = idx.copy()
        b = lambda q: (idx >> q) & 1  # noqa: E731
        if op == OP["CNOT"]:
            perm = np.where(b(q0) == 1, idx ^ (1 << q1), idx)
        elif op == OP["TOFFOLI"]:
            perm = np.where((b(q0) & b(q1)) == 1, idx ^ (1 << q2), idx)
        else:
            c = np.ones(N, dtype=bool) if op == OP["SWAP"] else b(q0) == 1
            t1, t2 = (q0, q1) if op == OP["SWAP"] else (q1, q2)
            diff = b(t1) != b(t2)
            perm = np.where(c & diff, idx ^ (1 << t1) ^ (1 << t2), idx)
        U = np.zeros((N, N), dtype=np.complex128)
        U[perm, idx] = 1.0
        return U
    if op in (OP["CZ"], OP["CPHASE"]):
        f = -1.0 if op == OP["CZ"] else np.exp(1j * th)
        both = (((idx >> q0) & 1) & ((idx >> q1) & 1)) == 1
        return np.diag(np.where(both, f, 1.0).astype(np.complex128))
    if op == OP["CY"]:
        U = np.zeros((N, N), dtype=np.complex128)
        for i in range(N):
            if (i >> q0) & 1:
                j = i ^ (1 << q1)
                bit = (i >> q1) & 1  # column i, input bit
                U[j, i] = 1j if bit == 0 else -1j
            else:
                U[i, i] = 1.0
        return U
    m = one_qubit_matrix(op, th)
    U = np.zeros((N, N), dtype=np.complex128)
    for i in range(N):
        bi = (i >> q0) & 1
        for bo in (0, 1):
            U[(i & ~(1 << q0)) | (bo << q0), i] = m[bo, bi]
    return U


def arity(op: int) -> int:
    if op in (OP["CNOT"], OP["CZ"], OP["CPHASE"], OP["SWAP"], OP["U4"], OP["CY"]):
        return 2
    if op in (OP["CSWAP"], OP["TOFFOLI"]):
        return 3
    return 1


def _apply_left(rho: np.ndarray, n: int, qs, m: np.ndarray) -> np.ndarray:
    """m (2^k x 2^k, basis index = bits of qs in order, qs[0] lowest) applied
    to the row index of rho, which has shape (2,)*2n (axis n-1-q = row bit q,
    axis 2n-1-q = column bit q)."""
    k = len(qs)
    t = m.reshape((2,) * (2 * k))  # out bits (high..low), in bits (high..low)
    axes = [n - 1 - q for q in reversed(qs)]  # matches the in-bit order of t
    out = np.tensordot(t, rho, axes=(list(range(k, 2 * k)), axes))
    return np.moveaxis(out, list(range(k)), axes)


def _apply_right_dagger(rho: np.ndarray, n: int, qs, m: np.ndarray) -> np.ndarray:
    """rho m^dagger: m^* applied to the column index."""
    k = len(qs)
    t = np.conj(m).reshape((2,) * (2 * k))
    axes = [2 * n - 1 - q for q in reversed(qs)]
    out = np.tensordot(t, rho, axes=(list(range(k, 2 * k)), axes))
    return np.moveaxis(out, list(range(k)), axes)


def _gate_local(g):
    """(qubits, matrix) of a gate, the matrix over the listed qubits."""
    op, q0, q1, q2, th = int(g["op"]), int(g["q0"]), int(g["q1"]), int(g["q2"]), float(g["theta"])
    ar = arity(op)
    qs = [q0, q1, q2][:ar]
    if ar == 1:
        return qs, one_qubit_matrix(op, th)
    return qs, full_unitary(ar, np.array((op, *range(ar), th, 0), dtype=g.dtype)
                            if ar == 3 else np.array((op, 0, 1, 0, th, 0), dtype=g.dtype))


def noisy_probabilities(n: int, gates, p1: float, p2: float, mistake: str = "") -> np.ndarray:
    """Z-basis outcome distribution of the trajectory ensemble.

    mistake: "" (the channel), "no_noise", "first_operand_only" (a multi-qubit
    gate's error only on its first operand), "before_gate" (errors drawn
    before the gate instead of after), "no_y" (X/Z only, each p/2),
    "p_swapped" (p2 after one-qubit gates and p1 after the others)."""
    N = 1 << n
    rho = np.zeros((N, N), dtype=np.complex128)
    rho[0, 0] = 1.0
    rho = rho.reshape((2,) * (2 * n))

    def conj_by(rho, q, P):
        return _apply_right_dagger(_apply_left(rho, n, [q], P), n, [q], P)

    def channel(rho, q, p):
        if p <= 0.0:
            return rho
        if mistake != "no_y":
            # the Pauli mixture in closed form: on qubit q's (row, column)
            # bits (a, b), diagonal blocks mix with their flipped partner,
            # off-diagonal blocks shrink by 1 - 4p/3
            r, c = n - 1 - q, 2 * n - 1 - q
            out = np.empty_like(rho)
            for a in (0, 1):
                for b in (0, 1):
                    sl = [slice(None)] * (2 * n)
                    sl[r], sl[c] = a, b
                    if a == b:
                        fl = list(sl)
                        fl[r], fl[c] = 1 - a, 1 - b
                        out[tuple(sl)] = (1 - 2 * p / 3) * rho[tuple(sl)] + \
                            (2 * p / 3) * rho[tuple(fl)]
                    else:
                        out[tuple(sl)] = (1 - 4 * p / 3) * rho[tuple(sl)]
            return out
        if mistake == "no_y":
            return (1 - p) * rho + p / 2 * (conj_by(rho, q, X) + conj_by(rho, q, Z))
        return (1 - p) * rho + p / 3 * (conj_by(rho, q, X) + conj_by(rho, q, Y) +
                                        conj_by(rho, q, Z))

    for g in gates:
        op = int(g["op"])
        ar = arity(op)
        p = p1 if ar == 1 else p2
        if mistake == "p_swapped":
            p = p2 if ar == 1 else p1
        qs, m = _gate_local(g)
        noisy_qs = qs[:1] if mistake == "first_operand_only" else qs
        if mistake == "before_gate":
            for q in noisy_qs:
                rho = channel(rho, q, p)
        rho = _apply_right_dagger(_apply_left(rho, n, qs, m), n, qs, m)
        if mistake not in ("before_gate", "no_noise"):
            for q in noisy_qs:
                rho = channel(rho, q, p)
    pr = np.real(np.diagonal(rho.reshape(N, N))).copy()
    pr[pr < 0] = 0.0
    return pr / pr.sum()


def marginal(probs: np.ndarray, n: int, qubits) -> np.ndarray:
    """Distribution of the bits `qubits` (qubits[j] -> bit j of the result)."""
    idx = np.arange(1 << n)
    key = np.zeros_like(idx)
    for j, q in enumerate(qubits):
        key |= ((idx >> q) & 1) << j
    return np.bincount(key, weights=probs, minlength=1 << len(qubits))


def marginal_of_samples(samples: np.ndarray, qubits) -> np.ndarray:
    s = np.asarray(samples, dtype=np.int64)
    key = np.zeros_like(s)
    for j, q in enumerate(qubits):
        key |= ((s >> q) & 1) << j
    return np.bincount(key, minlength=1 << len(qubits)) / s.size


def tvd(a: np.ndarray, b: np.ndarray) -> float:
    return 0.5 * float(np.abs(a - b).sum())


def tvd_bound(p: np.ndarray, samples: int, sigmas: float = 5.0) -> float:
    """Acceptance bound for the TVD between an empirical histogram of
    `samples` draws and the distribution p it was drawn from: the mean of
    the multinomial TVD, at most 1/2 sum_i sqrt(p_i (1 - p_i) / N), plus
    `sigmas` standard deviations of its bounded-differences (McDiarmid)
    tail, where one draw moves the TVD by at most 1/N (sigma = 1/(2 sqrt N))."""
    mean = 0.5 * float(np.sum(np.sqrt(p * (1 - p) / samples)))
    return mean + sigmas / (2.0 * math.sqrt(samples))
