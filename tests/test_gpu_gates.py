"""HIP layer-1 sweeps against the oracle: every gate, every target, bit for bit."""
import math

import numpy as np
import pytest

import pyoracle as po
from kat_util import arr, assert_kat, kats, random_state
from qoa_amd import StateVector
from qoa_amd._native import GATE_DTYPE
from test_oracle_kats import run_l1_oracle

pytestmark = pytest.mark.gpu

L1 = [c for c in kats() if c["level"] == "l1"]


def run_l1_hip(sv, fn, args):
    if fn == "hadamard":
        sv.apply_hadamard(args[2], complex(args[0], args[1]))
    elif fn in ("x", "y", "z", "t", "s", "reset"):
        getattr(sv, "apply_" + fn)(args[0])
    elif fn == "phaseshift":
        sv.apply_phaseshift(args[0], args[1])
    elif fn in ("rx", "ry", "rz"):
        getattr(sv, "apply_" + fn)(args[0], args[1])
    elif fn == "swap":
        sv.apply_swap(args[0], args[1])
    elif fn == "controlled_swap":
        sv.apply_controlled_swap(*args)
    elif fn == "cnot":
        sv.apply_cnot(args[0], args[1])
    elif fn == "cz":
        sv.apply_cz(args[0], args[1])
    elif fn == "cphase":
        sv.apply_controlled_phase_rotation(args[0], args[1], args[2])
    elif fn == "reset_all":
        sv.apply_reset_all()
    else:
        raise KeyError(fn)


@pytest.mark.parametrize("case", L1, ids=[c["name"] for c in L1])
def test_hip_kat(case):
    inp = arr(case["input"])
    n = int(math.log2(inp.size))
    with StateVector(n) as sv:
        sv.upload(inp)
        run_l1_hip(sv, case["fn"], case["args"])
        out = sv.download()
    assert_kat(out, arr(case["expected"]), case["eps"], case["name"])
    ref = inp.copy()
    run_l1_oracle(case["fn"], case["args"], ref)
    assert np.array_equal(out, ref), case["name"]


ONEQ = [("hadamard", lambda q: [1 / math.sqrt(2), 0.0, q]), ("x", lambda q: [q]),
        ("y", lambda q: [q]), ("z", lambda q: [q]), ("t", lambda q: [q]), ("s", lambda q: [q]),
        ("phaseshift", lambda q: [q, 0.731]), ("rx", lambda q: [q, 1.234]),
        ("ry", lambda q: [q, -2.5]), ("rz", lambda q: [q, 0.4321]), ("reset", lambda q: [q])]


@pytest.mark.parametrize("n", [1, 3, 7, 12])
@pytest.mark.parametrize("name,mk", ONEQ, ids=[o[0] for o in ONEQ])
def test_one_qubit_all_targets(n, name, mk):
    """Targets 0..n-1 plus out-of-range n and n+3 (reference no-op / global-phase rules)."""
    a0 = random_state(n, seed=17 * n + len(name))
    with StateVector(n) as sv:
        for q in list(range(n)) + [n, n + 3]:
            ref = a0.copy()
            run_l1_oracle(name, mk(q), ref)
            sv.upload(a0)
            run_l1_hip(sv, name, mk(q))
            out = sv.download()
            if name == "reset" and n > 2:
                # the bit-0 norm is a reduction: its summation order differs
                # from the CPU loop at the ulp level, the collapse itself is exact
                assert np.array_equal(out == 0, ref == 0), f"{name} n={n} q={q}"
                assert np.max(np.abs(out - ref)) <= 4e-16, f"{name} n={n} q={q}"
            else:
                assert np.array_equal(out, ref), f"{name} n={n} q={q}"


@pytest.mark.parametrize("n", [2, 4, 8, 13])
def test_two_qubit_all_pairs(n):
    a0 = random_state(n, seed=99 + n)
    pairs = [(c, t) for c in range(n) for t in range(n)]
    if n > 8:
        pairs = pairs[:: max(1, len(pairs) // 40)]
    with StateVector(n) as sv:
        for c, t in pairs + [(0, n), (n, 0)]:
            for name, args in (("cnot", [c, t]), ("swap", [c, t]),
                               ("cz", [1 << c, 1 << t]), ("cphase", [1 << c, 1 << t, 0.3]),
                               ("cz", [c, t]), ("cphase", [c, t, 1.1])):
                if name in ("cz", "cphase") and (args[0] >= (1 << 40) or args[1] >= (1 << 40)):
                    continue
                ref = a0.copy()
                run_l1_oracle(name, args, ref)
                sv.upload(a0)
                run_l1_hip(sv, name, args)
                assert np.array_equal(sv.download(), ref), f"{name} {args} n={n}"


@pytest.mark.parametrize("n", [3, 6, 9])
def test_controlled_swap_and_extended(n):
    a0 = random_state(n, seed=5 + n)
    rng = np.random.default_rng(n)
    u2 = np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))[0]
    u4 = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]
    with StateVector(n) as sv:
        for c in range(n):
            for t1 in range(n):
                for t2 in range(n):
                    ref = a0.copy()
                    po.orc.controlled_swap(ref, c, t1, t2)
                    sv.upload(a0)
                    sv.apply_controlled_swap(c, t1, t2)
                    assert np.array_equal(sv.download(), ref), (c, t1, t2)
                    if len({c, t1, t2}) == 3:
                        ref = a0.copy()
                        po.orc.toffoli(ref, c, t1, t2)
                        sv.upload(a0)
                        sv.apply_toffoli(c, t1, t2)
                        assert np.array_equal(sv.download(), ref), ("toffoli", c, t1, t2)
            for t in range(n):
                if t != c:
                    ref = a0.copy()
                    po.orc.controlled_y(ref, c, t)
                    sv.upload(a0)
                    sv.apply_controlled_y(c, t)
                    assert np.array_equal(sv.download(), ref), ("cy", c, t)
                    ref = a0.copy()
                    po.orc.two_qubit_unitary(ref, c, t, np.ascontiguousarray(u4.reshape(16)))
                    sv.upload(a0)
                    sv.apply_matrix4(c, t, u4)
                    assert np.array_equal(sv.download(), ref), ("u4", c, t)
            ref = a0.copy()
            po.orc.matrix2(ref, c, np.ascontiguousarray(u2.reshape(4)))
            sv.upload(a0)
            sv.apply_matrix2(c, u2)
            assert np.array_equal(sv.download(), ref), ("u2", c)


@pytest.mark.parametrize("n", [4, 10])
def test_pauli_codes_fused_sweep(n):
    a0 = random_state(n, seed=3 * n)
    rng = np.random.default_rng(7)
    with StateVector(n) as sv:
        for _ in range(12):
            codes = rng.integers(0, 4, size=n).astype(np.uint8)
            ref = a0.copy()
            po.orc.pauli_codes(ref, codes, n)
            sv.upload(a0)
            sv.apply_pauli_codes(codes)
            assert np.array_equal(sv.download(), ref), codes


def test_copy_is_complete_before_the_source_changes():
    """qoa_state_copy reads the source on the destination's stream; work
    enqueued right after it on the source's own stream must not change what
    the copy sees (2 GiB states, so the copy takes long enough to overlap)."""
    from qoa_amd import workloads as W

    n = 27
    with StateVector(n) as a, StateVector(n) as b, StateVector(n) as c:
        a.run_gates(W.layered_random(n, 1, seed=3))
        c.copy_from(a)
        b.copy_from(a)
        a.apply_reset_all()  # overwrites the source at once
        d, f = b.compare(c)
        assert d == 0.0 and abs(f - 1.0) < 1e-12


def test_32_qubit_state_index_path():
    """A 32-qubit state (64 GiB): gates on qubits 30/31 and masks with bits
    >= 31 run the 64-bit index kernels at real size. Checked by device
    reductions only (norm, measure p0/p1, samples), against the same gates
    on a 3-qubit oracle state where qubits (0, 30, 31) become (0, 1, 2), and
    the fused run against the per-gate run on a second 64 GiB state."""
    from qoa_amd._native import QOA_ERR_OOM, QoaError
    from qoa_amd.server import sample_state

    try:
        sv = StateVector(32)
    except QoaError as e:
        if e.code == QOA_ERR_OOM:
            pytest.skip("device has less than 64 GiB free for a 32-qubit state")
        raise
    m = {0: 0, 30: 1, 31: 2}
    ref = po.state(3)
    b = lambda q: 1 << q  # noqa: E731
    with sv:
        steps = [
            ("hadamard", (31,)), ("cnot", (31, 0)), ("hadamard", (30,)), ("rx", (30, 0.9)),
            ("phaseshift", (31, 1.3)), ("cz", (b(31), b(30))), ("cphase", (b(31), b(0), 0.7)),
            ("ry", (0, 2.2)), ("swap", (31, 0)), ("hadamard", (31,)), ("rz", (30, -0.4)),
        ]
        for fn, args in steps:
            if fn == "hadamard":
                sv.apply_hadamard(args[0])
                po.orc.hadamard(ref, 1 / math.sqrt(2), 0.0, m[args[0]])
            elif fn in ("rx", "ry", "rz", "phaseshift"):
                getattr(sv, "apply_" + fn)(args[0], args[1])
                getattr(po.orc, fn)(ref, m[args[0]], args[1])
            elif fn == "cnot":
                sv.apply_cnot(*args)
                po.orc.cnot(ref, m[args[0]], m[args[1]])
            elif fn == "swap":
                sv.apply_swap(*args)
                po.orc.swap(ref, m[args[0]], m[args[1]])
            elif fn == "cz":
                sv.apply_cz(*args)
                po.orc.cz(ref, 1 << m[args[0].bit_length() - 1], 1 << m[args[1].bit_length() - 1])
            elif fn == "cphase":
                sv.apply_controlled_phase_rotation(*args)
                po.orc.cphase(ref, 1 << m[args[0].bit_length() - 1],
                              1 << m[args[1].bit_length() - 1], args[2])
        assert abs(sv.norm_sqr() - 1.0) < 1e-13
        pr = np.abs(ref) ** 2
        samples = sample_state(sv, seed=5, count=4000).astype(np.int64)
        keep = (1 << 0) | (1 << 30) | (1 << 31)
        assert np.all((samples & ~keep) == 0)
        small = (samples & 1) | (((samples >> 30) & 1) << 1) | (((samples >> 31) & 1) << 2)
        counts = np.bincount(small, minlength=8)
        for i in range(8):
            e = 4000 * pr[i]
            assert abs(counts[i] - e) <= 5 * math.sqrt(e * (1 - pr[i])) + 1, (i, counts[i], e)
        want_p0 = float(pr[[0, 1, 2, 3]].sum())  # qubit 31 (oracle bit 2) = 0
        assert 0.05 < want_p0 < 0.95
        o, p0, p1 = sv.measure_with(31, 0.999)  # r >= p0: outcome 1
        assert o == 1 and abs(p0 - want_p0) < 1e-13 and abs(p1 - (1 - want_p0)) < 1e-13
        sv.normalize()
        post = sample_state(sv, seed=6, count=500).astype(np.int64)
        assert np.all(((post >> 31) & 1) == 1) and np.all((post & ~keep) == 0)
        # fused (64-bit tile indexing) equals per-gate on a second 64 GiB state
        g = np.array([(op, q0, q1, 0, th, 0) for op, q0, q1, th in [
            (0x05, 31, 0, 0.0), (0x17, 31, 2, 0.0), (0x22, 30, 0, 0.4), (0x35, 29, 31, 0.3),
            (0x23, 1, 0, 1.1), (0x17, 0, 30, 0.0), (0x1E, 31, 1, 0.0), (0x05, 2, 0, 0.0),
            (0x0F, 31, 0, 0.8), (0x0B, 30, 3, 0.0)]], dtype=GATE_DTYPE)
        with StateVector(32) as other:
            other.copy_from(sv)
            sv.run_fused(g)
            other.run_gates(g)
            d, f = sv.compare(other)
        assert d <= 1e-12 and f >= 1 - 1e-12, (d, f)
