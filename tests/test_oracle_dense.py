"""Pins every oracle gate against an independent dense-matrix construction:
textbook 2x2 matrices, numpy Kronecker products and projector sums, applied
to random states by a plain matrix-vector product."""
import itertools
import math

import numpy as np
import pytest

import pyoracle as po
from kat_util import random_state

I2 = np.eye(2, dtype=complex)
P0 = np.diag([1, 0]).astype(complex)
P1 = np.diag([0, 1]).astype(complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.diag([1, -1]).astype(complex)
H = np.array([[1, 1], [1, -1]], dtype=complex) / math.sqrt(2)


def ph(t):
    return np.diag([1, np.exp(1j * t)])


def rx(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -1j * s], [-1j * s, c]])


def ry(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -s], [s, c]], dtype=complex)


def rz(t):
    return np.diag([np.exp(-1j * t / 2), np.exp(1j * t / 2)])


def kron_ops(n, ops: dict):
    """Full operator: qubit q's factor at bit q (qubit n-1 is the most
    significant Kronecker factor)."""
    m = np.ones((1, 1), dtype=complex)
    for q in reversed(range(n)):
        m = np.kron(m, ops.get(q, I2))
    return m


def controlled(n, controls, target, u):
    """sum over control patterns: all-ones -> u on target, otherwise identity."""
    full = np.zeros((1 << n, 1 << n), dtype=complex)
    for pattern in itertools.product([0, 1], repeat=len(controls)):
        ops = {c: (P1 if b else P0) for c, b in zip(controls, pattern)}
        if all(pattern):
            ops[target] = u
        full += kron_ops(n, ops)
    return full


def swap_full(n, a, b):
    return controlled(n, [a], b, X) @ controlled(n, [b], a, X) @ controlled(n, [a], b, X)


def check(n, oracle_call, full, seed=0):
    a = random_state(n, seed)
    expect = full @ a
    got = a.copy()
    oracle_call(got)
    assert np.max(np.abs(got - expect)) < 1e-13


ONE_Q = {
    "H": (lambda a, q: po.orc.hadamard(a, 1 / math.sqrt(2), 0.0, q), lambda th: H),
    "X": (lambda a, q: po.orc.x(a, q), lambda th: X),
    "Y": (lambda a, q: po.orc.y(a, q), lambda th: Y),
    "Z": (lambda a, q: po.orc.z(a, q), lambda th: Z),
    "S": (lambda a, q: po.orc.s(a, q), lambda th: ph(math.pi / 2)),
    "T": (lambda a, q: po.orc.t(a, q), lambda th: ph(math.pi / 4)),
    "PHASE": (lambda a, q, th=0.9: po.orc.phaseshift(a, q, th), lambda th: ph(0.9)),
    "RX": (lambda a, q: po.orc.rx(a, q, 0.77), lambda th: rx(0.77)),
    "RY": (lambda a, q: po.orc.ry(a, q, -1.3), lambda th: ry(-1.3)),
    "RZ": (lambda a, q: po.orc.rz(a, q, 2.2), lambda th: rz(2.2)),
}


@pytest.mark.parametrize("n", [1, 3, 5])
@pytest.mark.parametrize("name", list(ONE_Q))
def test_one_qubit_gates(n, name):
    call, mat = ONE_Q[name]
    for q in range(n):
        check(n, lambda a: call(a, q), kron_ops(n, {q: mat(0)}), seed=q)


@pytest.mark.parametrize("n", [2, 4, 5])
def test_two_qubit_gates(n):
    for c, t in itertools.permutations(range(n), 2):
        check(n, lambda a: po.orc.cnot(a, c, t), controlled(n, [c], t, X), seed=c * 7 + t)
        check(n, lambda a: po.orc.cz(a, 1 << c, 1 << t), controlled(n, [c], t, Z))
        check(n, lambda a: po.orc.cphase(a, 1 << c, 1 << t, 0.6), controlled(n, [c], t, ph(0.6)))
        check(n, lambda a: po.orc.swap(a, c, t), swap_full(n, c, t))
        check(n, lambda a: po.orc.controlled_y(a, c, t), controlled(n, [c], t, Y))


@pytest.mark.parametrize("n", [3, 5])
def test_three_qubit_gates(n):
    for c, t1, t2 in itertools.permutations(range(n), 3):
        cswap = controlled(n, [c], t1, I2)  # identity placeholder
        full = kron_ops(n, {c: P0}) + kron_ops(n, {c: P1}) @ swap_full(n, t1, t2)
        check(n, lambda a: po.orc.controlled_swap(a, c, t1, t2), full)
        check(n, lambda a: po.orc.toffoli(a, c, t1, t2), controlled(n, [c, t1], t2, X))
        del cswap


@pytest.mark.parametrize("n", [2, 4])
def test_general_unitaries(n):
    rng = np.random.default_rng(1)
    for q in range(n):
        u = np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))[0]
        check(n, lambda a: po.orc.matrix2(a, q, np.ascontiguousarray(u.reshape(4))),
              kron_ops(n, {q: u}))
    for q1, q2 in itertools.permutations(range(n), 2):
        u = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]
        lo, hi = min(q1, q2), max(q1, q2)
        # the 4x4 acts on the (high, low) pair: sum over its matrix elements
        full = np.zeros((1 << n, 1 << n), dtype=complex)
        basis = {0: P0, 1: P1}
        E = lambda r, s: np.outer(np.eye(2)[r], np.eye(2)[s]).astype(complex)
        for r in range(4):
            for s in range(4):
                full += u[r, s] * kron_ops(n, {hi: E(r >> 1, s >> 1), lo: E(r & 1, s & 1)})
        check(n, lambda a: po.orc.two_qubit_unitary(a, q1, q2, np.ascontiguousarray(u.reshape(16))),
              full)
        del basis


def test_reference_cz_mask_quirk():
    """QuantumState::apply_cz(c, t) hands indices to a mask-taking kernel
    (src/runtime/quantum_state.rs:338): CZ(0,1) acts as Z on qubit 0 and
    CZ(1,2) as CZ on qubits (0,1); CZ(3,4) as CCZ on qubits {0,1,2}."""
    n = 4
    check(n, lambda a: po.orc.cz(a, 0, 1), kron_ops(n, {0: Z}))
    check(n, lambda a: po.orc.cz(a, 1, 2), controlled(n, [1], 0, Z))
    check(n, lambda a: po.orc.cz(a, 3, 4), controlled(n, [0, 1], 2, Z))


@pytest.mark.parametrize("n", [1, 3, 5])
def test_reset_is_projection(n):
    for q in range(n):
        a = random_state(n, 50 + q)
        proj = kron_ops(n, {q: P0}) @ a
        expect = proj / np.linalg.norm(proj)
        got = a.copy()
        po.orc.reset(got, q)
        assert np.max(np.abs(got - expect)) < 1e-14


@pytest.mark.parametrize("n", [2, 4])
def test_measure_is_projection(n):
    for q in range(n):
        a = random_state(n, 80 + q)
        p0 = np.linalg.norm(kron_ops(n, {q: P0}) @ a) ** 2
        for r, out in ((p0 * 0.5, 0), (p0 + (1 - p0) * 0.5, 1)):
            proj = kron_ops(n, {q: P0 if out == 0 else P1}) @ a
            got = a.copy()
            o, gp0, gp1 = po.measure(got, q, r)
            assert o == out
            assert abs(gp0 - p0) < 1e-14 and abs(gp1 - (1 - p0)) < 1e-14
            assert np.max(np.abs(got - proj / np.linalg.norm(proj))) < 1e-14


def test_measure_zero_probability_branch():
    """factor 0 when p <= 1e-12 (quantum_state.rs:431-443): the state vanishes."""
    a = np.zeros(4, dtype=np.complex128)
    a[0] = 1.0
    o, p0, p1 = po.measure(a, 0, 1.5)  # r >= p0 forces outcome 1 with p1 = 0
    assert o == 1 and p1 == 0.0 and not a.any()
