"""Closed forms, invariants and RNG known answers that pin the oracle."""
import math

import numpy as np
import pytest

import pyoracle as po
from kat_util import random_state
from qoa_amd import workloads as W
from qoa_amd._native import GATE_DTYPE, NOISE_FIXED, OP


def test_hadamard_involution_and_rotation_inverse():
    a0 = random_state(6, 1)
    a = a0.copy()
    for q in range(6):
        po.orc.hadamard(a, 1 / math.sqrt(2), 0.0, q)
        po.orc.hadamard(a, 1 / math.sqrt(2), 0.0, q)
    assert np.max(np.abs(a - a0)) < 1e-14
    a = a0.copy()
    po.orc.rx(a, 3, 0.7)
    po.orc.rx(a, 3, -0.7)
    po.orc.ry(a, 1, 1.9)
    po.orc.ry(a, 1, -1.9)
    assert np.max(np.abs(a - a0)) < 1e-14


def test_rz_is_phase_times_global_factor():
    th = 1.234
    a0 = random_state(5, 2)
    a = a0.copy()
    po.orc.rz(a, 2, th)
    b = a0.copy()
    po.orc.phaseshift(b, 2, th)
    b *= np.exp(-1j * th / 2)
    assert np.max(np.abs(a - b)) < 1e-14


def test_s_constant_is_from_polar_not_exact_i():
    """apply_s_vectorized uses from_polar(1, pi/2) = (6.123e-17, 1.0)."""
    a = np.array([0, 1], dtype=np.complex128)
    po.orc.s(a, 0)
    assert a[1].real == math.cos(math.pi / 2) and a[1].real != 0.0 and a[1].imag == 1.0


def test_unitary_circuit_preserves_norm():
    g = W.layered_random(12, 10, seed=3, ladders=True)
    a = po.state(12)
    po.run_gates(a, g)
    assert abs(po.norm_sqr(a) - 1.0) < 1e-12


@pytest.mark.parametrize("n", [3, 8, 12])
def test_permutation_gates_are_exact_index_maps(n):
    """X, CNOT and SWAP move amplitudes without arithmetic: compare the
    amplitude at every index with the closed-form image index (integer work,
    bit-exact)."""
    a0 = random_state(n, 9)
    idx = np.arange(1 << n)
    for q in range(n):
        a = a0.copy()
        po.orc.x(a, q)
        assert np.array_equal(a, a0[idx ^ (1 << q)])
    for c in range(n):
        for t in range(n):
            if c == t:
                continue
            a = a0.copy()
            po.orc.cnot(a, c, t)
            src = np.where((idx >> c) & 1 == 1, idx ^ (1 << t), idx)
            assert np.array_equal(a, a0[src])
            a = a0.copy()
            po.orc.swap(a, c, t)
            bc, bt = (idx >> c) & 1, (idx >> t) & 1
            src = idx ^ ((bc ^ bt) << c) ^ ((bc ^ bt) << t)
            assert np.array_equal(a, a0[src])


def test_diagonal_gates_commute():
    a0 = random_state(7, 4)
    a = a0.copy()
    po.orc.t(a, 2)
    po.orc.cphase(a, 1 << 3, 1 << 5, 0.4)
    po.orc.z(a, 6)
    b = a0.copy()
    po.orc.z(b, 6)
    po.orc.cphase(b, 1 << 3, 1 << 5, 0.4)
    po.orc.t(b, 2)
    assert np.max(np.abs(a - b)) < 1e-15


def test_ghz_and_qft3_closed_forms():
    n = 5
    a = po.state(n)
    po.orc.hadamard(a, 1 / math.sqrt(2), 0.0, 0)
    for q in range(1, n):
        po.orc.cnot(a, 0, q)
    expect = np.zeros(1 << n, dtype=complex)
    expect[0] = expect[-1] = 1 / math.sqrt(2)
    assert np.max(np.abs(a - expect)) < 1e-15
    # 3-qubit QFT of |1> (textbook circuit: H, controlled phases, final swap)
    a = np.zeros(8, dtype=np.complex128)
    a[1] = 1.0
    po.orc.hadamard(a, 1 / math.sqrt(2), 0.0, 2)
    po.orc.cphase(a, 1 << 1, 1 << 2, math.pi / 2)
    po.orc.cphase(a, 1 << 0, 1 << 2, math.pi / 4)
    po.orc.hadamard(a, 1 / math.sqrt(2), 0.0, 1)
    po.orc.cphase(a, 1 << 0, 1 << 1, math.pi / 2)
    po.orc.hadamard(a, 1 / math.sqrt(2), 0.0, 0)
    po.orc.swap(a, 0, 2)
    expect = np.exp(2j * math.pi * np.arange(8) / 8) / math.sqrt(8)
    assert np.max(np.abs(a - expect)) < 1e-15


def test_normalize_semantics():
    a = random_state(6, 5, normalise=False)
    s = np.vdot(a, a).real
    b = a.copy()
    assert po.orc.normalize(b) == 0
    assert np.max(np.abs(b - a / math.sqrt(s))) < 1e-15
    z = np.zeros(8, dtype=np.complex128)
    z[3] = 1e-7  # |a|^2 = 1e-14 <= 1e-12: reset to |0...0>
    assert po.orc.normalize(z) == 1
    assert z[0] == 1 and np.count_nonzero(z) == 1


def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32-10."""
    assert po.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert po.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                            0x6D5451FD]
    assert po.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) \
        == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_splitmix64_known_answers():
    assert po.splitmix64(1234567, 5) == [6457827717110365317, 3203168211198807973,
                                         9817491932198370423, 4593380528125082431,
                                         16408922859458223821]


def test_depolarizing_channel_statistics():
    """Fixed(p) after one gate on n qubits: each qubit takes an error with
    probability p, X/Y/Z equally often (quantum_state.rs:872-885)."""
    n, p, trials = 3, 0.3, 4000
    errs = np.zeros(4)
    g = np.array([(OP["Z"], 0, 0, 0, 0.0, 0)], dtype=GATE_DTYPE)
    for t in range(trials):
        qs = po.QS(n, NOISE_FIXED, p, seed=1000 + t)
        assert qs.apply(g) == 0
        k = int(np.argmax(np.abs(qs.a)))
        errs[bin(k).count("1")] += 1
    # |000> stays unless an X or Y hits a qubit: P(flip) = 2p/3 per qubit
    pf = 2 * p / 3
    expect = [math.comb(n, j) * pf ** j * (1 - pf) ** (n - j) * trials for j in range(n + 1)]
    for o, e in zip(errs, expect):
        assert abs(o - e) < 5 * math.sqrt(e) + 5


@pytest.mark.parametrize("n", [1, 5, 13, 14])
def test_sampler_inverts_the_cdf(n):
    """orc_sample returns the index whose CDF interval holds u*total (checked
    against numpy's cumulative sum away from interval boundaries), never an
    index of probability zero, and the last nonzero index for u -> 1."""
    a = random_state(n, seed=n)
    a[::3] = 0  # zero-probability entries are never chosen
    p = np.abs(a) ** 2
    cdf = np.cumsum(p)
    L = po.lib()
    rng = np.random.default_rng(n)
    for u in rng.random(200):
        t = u * cdf[-1]
        want = int(np.searchsorted(cdf, t, side="right"))
        got = L.orc_sample(po._p(a), a.size, float(u))
        if np.min(np.abs(cdf - t)) > 1e-12:
            assert got == want, (u, got, want)
        assert p[got] > 0
    last_nz = int(np.nonzero(p)[0][-1])
    assert L.orc_sample(po._p(a), a.size, 1.0 - 2.0 ** -53) == last_nz
    assert L.orc_sample(po._p(a), a.size, 0.0) == int(np.nonzero(p)[0][0])


def test_sample_uniforms_are_distinct_streams():
    """The sampler's uniforms lie in [0,1), differ across shots, trajectories
    and seeds, and have the mean and variance of U(0,1)."""
    L = po.lib()
    u = np.array([L.orc_sample_uniform(5, 0, s) for s in range(4000)])
    assert np.all((u >= 0) & (u < 1)) and len(set(u.tolist())) == u.size
    assert abs(u.mean() - 0.5) < 0.02 and abs(u.var() - 1 / 12) < 0.01
    assert L.orc_sample_uniform(5, 1, 0) != L.orc_sample_uniform(5, 0, 0)
    assert L.orc_sample_uniform(6, 0, 0) != L.orc_sample_uniform(5, 0, 0)


def test_threaded_gate_loops_are_the_same_arithmetic():
    """The all-core CPU baseline splits the gate loops over threads; every
    iteration owns its amplitudes, so results stay bit-identical. The
    faithful baseline (sweep + normalize per gate) equals the sequential run
    normalised, to rounding."""
    n = 12
    g = W.layered_random(n, 10, seed=7, ladders=True)
    a = random_state(n, seed=7)
    seq, par, fai = a.copy(), a.copy(), a.copy()
    po.run_gates(seq, g)
    po.set_threads(4)
    try:
        po.run_gates(par, g)
    finally:
        po.set_threads(1)
    assert np.array_equal(seq, par)
    po.run_gates_faithful(fai, g, 4)
    assert np.max(np.abs(fai - seq / np.linalg.norm(seq))) <= 1e-14
