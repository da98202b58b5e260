"""Dense 5-qubit blocks on FP32 matrix cores (complex64 states) against the
FP64 oracle: 1 - F <= 1e-5 and max |d amp| <= 1e-5 (the FP32 tolerance),
renormalised to unit norm; plus the device state comparison."""
import numpy as np
import pytest

import pyoracle as po
from kat_util import random_state
from qoa_amd import StateVector
from qoa_amd import _native
from qoa_amd import workloads as W
from qoa_amd._native import QOA_BLOCK_NO_RENORM, QOA_C64
from test_planner import mixed_circuit

pytestmark = pytest.mark.gpu


def fidelity(a, b):
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


@pytest.mark.parametrize("n", [5, 6, 8, 9, 10, 12, 15])
def test_blocks_layered_circuit_vs_oracle(n):
    g = W.layered_random(n, 20, seed=n, ladders=True)
    a0 = random_state(n, seed=n)
    ref = a0.copy()
    po.run_gates(ref, g)
    with StateVector(n, precision=QOA_C64) as sv:
        sv.upload(a0)
        nb = sv.run_blocks(g)
        out = sv.download()
    assert 0 < nb < len(g)
    assert 1 - fidelity(out, ref) <= 1e-5
    assert np.max(np.abs(out - ref)) <= 1e-5
    assert abs(np.vdot(out, out).real - 1) <= 1e-6


@pytest.mark.parametrize("seed", [1, 2])
def test_blocks_full_gate_surface(seed):
    n = 11
    g, mats = mixed_circuit(n, 400, seed=seed)
    a0 = random_state(n, seed=seed)
    ref = a0.copy()
    po.run_gates(ref, g, matrices=mats)
    with StateVector(n, precision=QOA_C64) as sv:
        sv.upload(a0)
        sv.run_blocks(g, matrices=mats, flags=QOA_BLOCK_NO_RENORM)
        out = sv.download()
    assert 1 - fidelity(out, ref) <= 1e-5
    assert np.max(np.abs(out - ref)) <= 1e-5


def test_single_block_is_the_matrix_product():
    """One block (all gates on qubits 0..4 of 10): exactly U applied to each
    column, to FP32 rounding."""
    n = 10
    g = W.layered_random(5, 6, seed=4)
    from qoa_amd import plan_blocks

    qs, us = plan_blocks(n, g)
    assert len(qs) == 1 and list(qs[0]) == [0, 1, 2, 3, 4]
    a0 = random_state(n, seed=3)
    want = (a0.reshape(-1, 32) @ us[0].T).reshape(-1)
    with StateVector(n, precision=QOA_C64) as sv:
        sv.upload(a0)
        sv.run_blocks(g, flags=QOA_BLOCK_NO_RENORM)
        out = sv.download()
    assert np.max(np.abs(out - want)) <= 2e-6


def test_compare_on_device():
    n = 12
    a = random_state(n, seed=1)
    b = a * np.exp(0.3j) + 1e-7 * random_state(n, seed=2)
    with StateVector(n) as x, StateVector(n) as y, StateVector(n, precision=QOA_C64) as z:
        x.upload(a)
        y.upload(b)
        z.upload(a)
        d, f = x.compare(y)
        assert abs(d - np.max(np.abs(a - b))) <= 1e-15
        assert abs(f - fidelity(a, b)) <= 1e-12
        d, f = x.compare(x)
        assert d == 0.0 and abs(f - 1) <= 1e-15
        d, f = x.compare(z)  # complex128 against its complex64 rounding
        assert 0 < d <= 1e-7 and 1 - f <= 1e-12


def test_c3_shape_against_per_gate_fp64():
    """The C3 comparison at 20 qubits: FP32 blocks against the unfused FP64
    device run, both resident, compared on the device."""
    n = 20
    g = W.layered_random(n, 30, seed=11, ladders=True)
    with StateVector(n) as ref, StateVector(n, precision=QOA_C64) as s32:
        ref.run_gates(g)
        s32.run_blocks(g)
        d, f = ref.compare(s32)
    assert 1 - f <= 1e-5 and d <= 1e-5


def test_blocks_index64_identical():
    lib = _native.load()
    n = 13
    g = W.layered_random(n, 10, seed=5, ladders=True)
    outs = []
    try:
        for force in (0, 1):
            _native.check(lib.qoa_set_option(b"index64", force))
            with StateVector(n, precision=QOA_C64) as sv:
                sv.run_blocks(g)
                outs.append(sv.download())
    finally:
        lib.qoa_set_option(b"index64", 0)
    assert np.array_equal(outs[0], outs[1])
