"""The dense-block fuser (host code, no GPU): the product of its 32x32
blocks, applied with numpy, reproduces the oracle's FP64 run of the circuit;
blocks are unitary and act on five distinct qubits."""
import numpy as np
import pytest

import pyoracle as po
from kat_util import random_state
from qoa_amd import QoaError, plan_blocks
from qoa_amd import workloads as W
from qoa_amd._native import GATE_DTYPE, OP, QOA_CZ_REF_MASKS
from test_planner import mixed_circuit


def apply_blocks(a: np.ndarray, n: int, qs: np.ndarray, us: np.ndarray) -> None:
    """Dense application: for every assignment of the other bits, the 32
    amplitudes indexed by the block qubits (local bit b = qs[b]) get U."""
    idx = np.arange(1 << n)
    for q, u in zip(qs, us):
        off = np.zeros(32, dtype=np.int64)
        for r in range(32):
            for b, qq in enumerate(q):
                if (r >> b) & 1:
                    off[r] |= 1 << int(qq)
        mask = sum(1 << int(qq) for qq in q)
        bases = idx[(idx & mask) == 0]
        sel = bases[:, None] | off[None, :]
        a[sel] = a[sel] @ u.T


@pytest.mark.parametrize("n,depth", [(5, 6), (7, 10), (10, 20), (12, 15)])
def test_block_product_matches_oracle(n, depth):
    g = W.layered_random(n, depth, seed=n, ladders=True)
    qs, us = plan_blocks(n, g)
    a = random_state(n, seed=n)
    ref = a.copy()
    po.run_gates(ref, g)
    apply_blocks(a, n, qs, us)
    assert np.max(np.abs(a - ref)) <= 1e-12
    assert len(qs) < len(g)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_block_product_full_gate_surface(seed):
    n = 9
    g, mats = mixed_circuit(n, 300, seed=seed)
    qs, us = plan_blocks(n, g, matrices=mats)
    a = random_state(n, seed=seed)
    ref = a.copy()
    po.run_gates(ref, g, matrices=mats)
    apply_blocks(a, n, qs, us)
    assert np.max(np.abs(a - ref)) <= 1e-12


def test_ref_mask_semantics_and_block_width():
    """QOA_CZ_REF_MASKS reads CZ(c, t) as the mask c|t; block_qubits < 5
    narrows the blocks (padded to five qubits with the identity)."""
    n = 8
    rows = [(OP["H"], q, 0, 0, 0.0, 0) for q in range(n)]
    rows += [(OP["CZ"], 3, 4, 0, 0.0, 0), (OP["CPHASE"], 1, 6, 0, 0.7, 0),
             (OP["CNOT"], 2, 5, 0, 0.0, 0)]
    g = np.array(rows, dtype=GATE_DTYPE)
    for flags, widths in ((0, (2, 3, 5)), (QOA_CZ_REF_MASKS, (3, 5))):  # masks 3|4, 1|6: 3 bits
        for bq in widths:
            qs, us = plan_blocks(n, g, block_qubits=bq, flags=flags)
            a = random_state(n, seed=bq)
            ref = a.copy()
            po.run_gates(ref, g, flags=flags)
            apply_blocks(a, n, qs, us)
            assert np.max(np.abs(a - ref)) <= 1e-12


def test_blocks_are_unitary_on_distinct_qubits():
    n = 14
    g = W.layered_random(n, 12, seed=9, ladders=True)
    qs, us = plan_blocks(n, g)
    eye = np.eye(32)
    for q, u in zip(qs, us):
        assert len(set(q.tolist())) == 5 and list(q) == sorted(q) and q.max() < n
        assert np.max(np.abs(u @ u.conj().T - eye)) <= 1e-12


def test_fuser_rejects_what_it_cannot_fuse():
    n = 8
    with pytest.raises(QoaError):  # not unitary
        plan_blocks(n, np.array([(OP["MEASURE"], 0, 0, 0, 0.0, 0)], dtype=GATE_DTYPE))
    with pytest.raises(QoaError):  # fewer than 5 qubits
        plan_blocks(4, np.array([(OP["H"], 0, 0, 0, 0.0, 0)], dtype=GATE_DTYPE))
    with pytest.raises(QoaError):  # REF mask 3|252 touches 8 qubits
        plan_blocks(n, np.array([(OP["CZ"], 3, 252, 0, 0.0, 0)], dtype=GATE_DTYPE),
                    flags=QOA_CZ_REF_MASKS)
    with pytest.raises(QoaError):  # block wider than 5 qubits
        plan_blocks(n, np.array([(OP["H"], 0, 0, 0, 0.0, 0)], dtype=GATE_DTYPE), block_qubits=6)
