"""The oracle against the reference's own known-answer tests (src/test.rs)."""
import numpy as np
import pytest

import pyoracle as po
from kat_util import arr, assert_kat, kats
from qoa_amd._native import GATE_DTYPE, OP

L1 = [c for c in kats() if c["level"] == "l1"]
QS = [c for c in kats() if c["level"] == "qs"]


def run_l1_oracle(fn, args, a):
    o = po.orc
    if fn == "hadamard":
        o.hadamard(a, args[0], args[1], args[2])
    elif fn in ("x", "y", "z", "t", "s", "reset"):
        getattr(o, fn)(a, args[0])
    elif fn == "phaseshift":
        o.phaseshift(a, args[0], args[1])
    elif fn in ("rx", "ry", "rz"):
        getattr(o, fn)(a, args[0], args[1])
    elif fn == "swap":
        o.swap(a, args[0], args[1])
    elif fn == "controlled_swap":
        o.controlled_swap(a, args[0], args[1], args[2])
    elif fn == "cnot":
        o.cnot(a, args[0], args[1])
    elif fn == "cz":
        o.cz(a, args[0], args[1])
    elif fn == "cphase":
        o.cphase(a, args[0], args[1], args[2])
    elif fn == "reset_all":
        o.reset_all(a)
    else:
        raise KeyError(fn)


@pytest.mark.parametrize("case", L1, ids=[c["name"] for c in L1])
def test_oracle_l1_kat(case):
    a = arr(case["input"])
    run_l1_oracle(case["fn"], case["args"], a)
    assert_kat(a, arr(case["expected"]), case["eps"], case["name"])


def qs_gates(fn, args):
    """execute_arithmetic's mnemonic -> gate rows (src/runtime/quantum_state.rs:460-560)."""
    if fn == "EBELL":
        return [(OP["H"], args[0], 0, 0, 0.0, 0), (OP["CNOT"], args[0], args[1], 0, 0.0, 0)]
    if fn == "RESET_ALL":
        return [(OP["RESET_ALL"], 0, 0, 0, 0.0, 0)]
    op = OP[fn]
    q = list(args) + [0, 0, 0]
    theta = 0.0
    if fn in ("PHASESHIFT",):
        q = [args[0], 0, 0]
        theta = args[1]
    return [(op, q[0], q[1], q[2], theta, 0)]


@pytest.mark.parametrize("case", QS, ids=[c["name"] for c in QS])
def test_oracle_qs_kat(case):
    inp = arr(case["input"])
    n = int(np.log2(inp.size))
    if case["fn"] == "QINIT":
        qs = po.QS(case["args"][0])  # QuantumState::new(max(n,1))
    else:
        qs = po.QS(n)
        qs.a[:] = inp
        for row in qs_gates(case["fn"], case["args"]):
            g = np.array([row], dtype=GATE_DTYPE)[0]
            assert qs.apply(np.array([row], dtype=GATE_DTYPE)) == 0
    assert_kat(qs.a, arr(case["expected"]), case["eps"], case["name"])
