"""Whole circuits, reductions and QuantumState semantics: HIP vs oracle."""
import numpy as np
import pytest

import pyoracle as po
from kat_util import random_state
from qoa_amd import QuantumState, StateVector
from qoa_amd import workloads as W
from qoa_amd._native import GATE_DTYPE, NOISE_FIXED, NOISE_RANDOM, OP, QOA_CZ_REF_MASKS

pytestmark = pytest.mark.gpu


def fidelity(a, b):
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


def test_c1_per_gate_bit_exact():
    """16-qubit, 20-layer random circuit (470 gates): one sweep per gate,
    identical to the CPU path bit for bit."""
    g = W.layered_random(16, 20, seed=1)
    ref = po.state(16)
    po.run_gates(ref, g)
    with StateVector(16) as sv:
        sv.run_gates(g)
        out = sv.download()
        probs = sv.probabilities()
    assert np.array_equal(out, ref)
    assert np.array_equal(probs, po.probabilities(ref))


def test_c2_shape_at_16_qubits_with_ladders_and_ref_masks():
    g = W.layered_random(16, 10, seed=2, ladders=True)
    for flags in (0, QOA_CZ_REF_MASKS):
        ref = po.state(16)
        po.run_gates(ref, g, flags=flags)
        with StateVector(16) as sv:
            sv.run_gates(g, flags=flags)
            assert np.array_equal(sv.download(), ref), flags


@pytest.mark.parametrize("n", [1, 5, 12, 20])
def test_norm_and_normalize(n):
    a = random_state(n, seed=n, normalise=False) * 3.7
    with StateVector(n) as sv:
        sv.upload(a)
        s = sv.norm_sqr()
        assert abs(s - po.norm_sqr(a)) <= 1e-13 * s
        s2, reset = sv.normalize()
        assert not reset and abs(s2 - s) <= 1e-13 * s
        ref = a.copy()
        po.orc.normalize(ref)
        assert np.max(np.abs(sv.download() - ref)) <= 1e-15


def test_normalize_single_amplitude_bit_exact():
    """With one non-zero amplitude the sum has one term, so sqrt and the
    division on the device must agree with the CPU bit for bit."""
    rng = np.random.default_rng(0)
    with StateVector(6) as sv:
        for _ in range(50):
            a = np.zeros(64, dtype=np.complex128)
            a[rng.integers(64)] = complex(*rng.standard_normal(2) * 10 ** rng.uniform(-5, 5))
            sv.upload(a)
            sv.normalize()
            ref = a.copy()
            po.orc.normalize(ref)
            assert np.array_equal(sv.download(), ref)


def test_normalize_zero_state_resets():
    with StateVector(5) as sv:
        sv.upload(np.zeros(32, dtype=np.complex128))
        s, reset = sv.normalize()
        assert reset and s == 0.0
        out = sv.download()
        assert out[0] == 1 and np.count_nonzero(out) == 1


@pytest.mark.parametrize("n", [3, 9, 16])
def test_measure_injected_uniform(n):
    a0 = random_state(n, seed=40 + n)
    with StateVector(n) as sv:
        for q in range(n):
            for r in (0.05, 0.5, 0.95):
                ref = a0.copy()
                o_ref, p0_ref, p1_ref = po.measure(ref, q, r)
                if abs(r - p0_ref) < 1e-9:
                    continue
                sv.upload(a0)
                o, p0, p1 = sv.measure_with(q, r)
                assert o == o_ref
                assert abs(p0 - p0_ref) < 1e-14 and abs(p1 - p1_ref) < 1e-14
                assert np.max(np.abs(sv.download() - ref)) < 1e-14


def test_measure_collapse_bit_exact_on_basis_states():
    """Collapse index predicate: exact on computational-basis superpositions."""
    n = 6
    with StateVector(n) as sv:
        for q in range(n):
            a = np.zeros(64, dtype=np.complex128)
            a[0] = 0.6
            a[(1 << q) | 5] = 0.8j
            for r in (0.1, 0.9):
                ref = a.copy()
                o_ref, _, _ = po.measure(ref, q, r)
                sv.upload(a)
                o, _, _ = sv.measure_with(q, r)
                assert o == o_ref
                assert np.array_equal(sv.download(), ref)


@pytest.mark.parametrize("kind,p", [(0, 0.0), (NOISE_FIXED, 0.3), (NOISE_RANDOM, 0.0)])
def test_quantum_state_sequence_with_noise(kind, p):
    """QuantumState wrappers: gate, apply_noise (same Philox draws), normalize."""
    n = 7
    g = W.layered_random(n, 6, seed=11)
    meas = [2, 5, 0]
    from qoa_amd import NoiseConfig

    nc = None if kind == 0 else (NoiseConfig.Fixed(p) if kind == NOISE_FIXED else NoiseConfig.Random())
    qs = QuantumState(n, nc)
    qs.seed_rng(1234)
    ref = po.QS(n, kind, p, seed=1234)
    outs, outs_ref = [], []
    for k, row in enumerate(g):
        one = np.array([row], dtype=GATE_DTYPE)
        qs._apply(int(row["op"]), int(row["q0"]), int(row["q1"]), int(row["q2"]), float(row["theta"]))
        assert ref.apply(one) == 0
        if k % 25 == 24:
            q = meas[(k // 25) % len(meas)]
            outs.append(qs.measure(q))
            outs_ref.append(ref.measure(q))
    assert outs == outs_ref
    out = qs.download()
    assert np.max(np.abs(out - ref.a)) <= 1e-12
    assert fidelity(out, ref.a) >= 1 - 1e-12
    qs.close()


def test_quantum_state_bounds_are_noops(capsys):
    qs = QuantumState(3)
    qs.apply_h(0)
    before = qs.download()
    qs.apply_h(3)
    qs.apply_cnot(1, 1)
    qs.apply_cz(0, 7)
    assert np.array_equal(qs.download(), before)
    err = capsys.readouterr().err
    assert "out of bounds" in err
    with pytest.raises(ValueError):
        qs.measure(3)
    qs.close()


def test_measure_all_leaves_state():
    qs = QuantumState(5)
    qs.seed_rng(9)
    for q in range(5):
        qs.apply_h(q)
    before = qs.download()
    bits = qs.measure_all()
    assert len(bits) == 5 and set(bits) <= {0, 1}
    assert np.array_equal(qs.download(), before)
    qs.close()


def test_validate_state_on_device():
    """validate_state (quantum_state.rs:135-167) with the checks on the device:
    NaN before Inf before the 1e-9 norm tolerance, messages as the reference."""
    from qoa_amd import QuantumState

    n = 12
    with QuantumState(n) as qs:
        qs.validate_state()
        a = random_state(n, seed=1)
        for bad, msg in ((np.nan, "nan"), (np.inf, "infinite")):
            b = a.copy()
            b[1234] = complex(bad, 0.0) if msg == "nan" else complex(0.0, bad)
            b[99] = complex(np.inf, 0.0) if msg == "nan" else b[99]  # NaN is reported first
            qs.upload(b)
            with pytest.raises(ValueError, match=msg):
                qs.validate_state()
        qs.upload(a * 1.001)
        with pytest.raises(ValueError, match="not normalized"):
            qs.validate_state()
        qs.upload(a)
        qs.validate_state()


def test_state_json_device_round_trip(tmp_path):
    """F2 (src/main.rs:3800-3827, 4290-4296): a device state saved as the serde
    document and loaded back into a new device state is bit-identical, with
    its noise configuration and registers."""
    from qoa_amd import NoiseConfig, stateio

    n = 10
    with QuantumState(n, NoiseConfig.Fixed(0.02)) as qs:
        qs.run_gates(W.layered_random(n, 6, seed=11, ladders=True))
        before = qs.download()
        path = tmp_path / "state.json"
        stateio.save_state(str(path), qs, registers=[1.5, -2.0])
    qs2, regs, meta = stateio.load_state(str(path))
    with qs2:
        assert qs2.n == n and regs == [1.5, -2.0]
        assert qs2.get_noise_config().kind == "Fixed" and qs2.get_noise_config().p == 0.02
        assert np.array_equal(qs2.download(), before)
