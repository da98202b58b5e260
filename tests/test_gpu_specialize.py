"""Specialised passes (QOA_FUSE_SPECIALIZE): the per-pass kernels compiled at
run time give the op-loop kernel's results bit for bit, for the exact, FMA
and merged-diagonal arithmetic, every gate kind, several tile sizes and the
64-bit index kernels; a second run of the same circuit structure reuses the
compiled code."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
from kat_util import random_state
from qoa_amd import StateVector
from qoa_amd import _native as N
from qoa_amd import workloads as W
from qoa_amd._native import (QOA_FUSE_FACTOR, QOA_FUSE_FMA, QOA_FUSE_MERGE_DIAG, QOA_FUSE_RELABEL,
                             QOA_FUSE_SPECIALIZE)
from test_planner import mixed_circuit

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _private_code_cache(tmp_path, monkeypatch):
    monkeypatch.setenv("QOA_SPEC_CACHE", str(tmp_path / "spec"))


def _compiles():
    c = C.c_uint32()
    assert N.load().qoa_specialize_info(None, C.byref(c), None) == 0
    return c.value


def _both(n, g, flags, tile=0, mats=None, a0=None):
    outs = []
    with StateVector(n) as sv:
        for f in (flags, flags | QOA_FUSE_SPECIALIZE):
            if a0 is None:
                sv.apply_reset_all()
            else:
                sv.upload(a0)
            sv.run_fused(g, tile_qubits=tile, flags=f, matrices=mats)
            outs.append(sv.download())
    return outs


@pytest.mark.parametrize("flags", [0, QOA_FUSE_FMA, QOA_FUSE_MERGE_DIAG, QOA_FUSE_FACTOR,
                                   QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG, QOA_FUSE_RELABEL,
                                   QOA_FUSE_RELABEL | QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG])
@pytest.mark.parametrize("n,tile", [(16, 0), (18, 12), (15, 10)])
def test_layered_circuit_equals_op_loop(n, tile, flags):
    g = W.layered_random(n, 12, seed=n + tile, ladders=True)
    loop, spec = _both(n, g, flags, tile)
    assert np.array_equal(loop, spec)
    ref = po.state(n)
    po.run_gates(ref, g)
    assert np.max(np.abs(spec - ref)) <= 1e-12


@pytest.mark.parametrize("flags", [0, QOA_FUSE_FMA, QOA_FUSE_FACTOR, QOA_FUSE_RELABEL])
@pytest.mark.parametrize("n,tile", [(12, 8), (14, 12)])
def test_mixed_surface_equals_op_loop(n, tile, flags):
    """U2/U4, Y, SWAP, CSWAP, Toffoli, CY, controls on register and thread
    bits and outside the tile."""
    g, mats = mixed_circuit(n, 300, seed=n * 5 + tile)
    a0 = random_state(n, seed=n + 1)
    loop, spec = _both(n, g, flags, tile, mats, a0)
    assert np.array_equal(loop, spec)


def test_index64_kernels_equal_op_loop():
    lib = N.load()
    g = W.layered_random(14, 8, seed=9, ladders=True)
    assert lib.qoa_set_option(b"index64", 1) == 0
    try:
        loop, spec = _both(14, g, QOA_FUSE_MERGE_DIAG)
    finally:
        lib.qoa_set_option(b"index64", 0)
    assert np.array_equal(loop, spec)


def test_same_structure_compiles_once():
    n = 14
    g = W.layered_random(n, 6, seed=77, ladders=True)
    with StateVector(n) as a, StateVector(n) as b:
        a.run_fused(g, flags=QOA_FUSE_MERGE_DIAG | QOA_FUSE_SPECIALIZE)
        c1 = _compiles()
        # other angles, same structure: new coefficients, the same kernels
        g2 = g.copy()
        g2["theta"] = g2["theta"] * 0.5 + 0.25
        b.run_fused(g2, flags=QOA_FUSE_MERGE_DIAG | QOA_FUSE_SPECIALIZE)
        assert _compiles() == c1
        out = b.download()
    ref = po.state(n)
    po.run_gates(ref, g2)
    assert np.max(np.abs(out - ref)) <= 1e-12
