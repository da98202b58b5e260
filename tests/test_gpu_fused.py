"""Fused LDS passes: bit-identical to the per-gate run of the planned order,
within 1e-12 of the original order, at every tile size."""
import numpy as np
import pytest

import pyoracle as po
from kat_util import random_state
from qoa_amd import StateVector, plan_fused
from qoa_amd import _native
from qoa_amd import workloads as W
from qoa_amd._native import QOA_CZ_REF_MASKS

QOA_FUSE_LDS_ONLY = 2
QOA_FUSE_FMA = 4
QOA_FUSE_MERGE_DIAG = 16
QOA_FUSE_FACTOR = 128
QOA_FUSE_RELABEL = 256
MODES = [0, QOA_FUSE_LDS_ONLY]  # register-resident segments (default) / LDS-resident
from test_planner import mixed_circuit

pytestmark = pytest.mark.gpu


def fidelity(a, b):
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("tile,low", [(4, 2), (6, 3), (8, 3), (10, 3), (12, 3), (12, 4), (13, 3)])
def test_c1_fused_matches_planned_order_bit_exact(tile, low, mode):
    n = 16
    g = W.layered_random(n, 20, seed=1)
    order, passes = plan_fused(n, g, tile_qubits=tile, low_qubits=low, flags=mode)
    ref = po.state(n)
    po.run_gates(ref, g, order=order)
    orig = po.state(n)
    po.run_gates(orig, g)
    with StateVector(n) as sv:
        sv.run_fused(g, tile_qubits=tile, low_qubits=low, flags=mode)
        out = sv.download()
    assert np.array_equal(out, ref), f"tile={tile} passes={passes}"
    assert np.max(np.abs(out - orig)) <= 1e-12
    assert fidelity(out, orig) >= 1 - 1e-12


@pytest.mark.parametrize("n,tile", [(10, 4), (10, 6), (12, 8), (14, 12), (15, 10)])
@pytest.mark.parametrize("flags", [0, QOA_CZ_REF_MASKS, QOA_FUSE_LDS_ONLY,
                                   QOA_FUSE_LDS_ONLY | QOA_CZ_REF_MASKS])
def test_mixed_gate_surface_fused(n, tile, flags):
    g, mats = mixed_circuit(n, 400, seed=n * 7 + tile)
    order, _ = plan_fused(n, g, tile_qubits=tile, flags=flags)
    a0 = random_state(n, seed=n)
    ref = a0.copy()
    po.run_gates(ref, g, flags=flags, matrices=mats, order=order)
    with StateVector(n) as sv:
        sv.upload(a0)
        sv.run_fused(g, tile_qubits=tile, flags=flags, matrices=mats)
        out = sv.download()
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("n", [1, 2, 5, 8, 9, 12, 13])
def test_whole_state_single_tile(n):
    g = W.layered_random(n, 12, seed=n, ladders=True)
    tile = 13 if n == 13 else 0  # the default tile holds 12 qubits
    order, passes = plan_fused(n, g, tile_qubits=tile)
    assert passes == 1
    ref = po.state(n)
    po.run_gates(ref, g, order=order)
    with StateVector(n) as sv:
        sv.run_fused(g, tile_qubits=tile)
        assert np.array_equal(sv.download(), ref)


def test_index64_paths_match():
    """The 64-bit index kernels (taken by states of 32+ qubits) give the same
    results as the 32-bit ones."""
    lib = _native.load()
    n = 14
    g = W.layered_random(n, 8, seed=3, ladders=True)
    gm, mats = mixed_circuit(n, 200, seed=5)
    outs = []
    try:
        for force in (0, 1):
            _native.check(lib.qoa_set_option(b"index64", force))
            with StateVector(n) as sv:
                sv.run_gates(g)
                sv.run_fused(g, tile_qubits=9)
                sv.run_gates(gm, matrices=mats)
                sv.run_fused(gm, tile_qubits=7, matrices=mats)
                sv.apply_reset(3)
                s, _ = sv.normalize()
                o, _, _ = sv.measure_with(5, 0.3)
                outs.append((sv.download(), sv.probabilities(), s, o))
    finally:
        lib.qoa_set_option(b"index64", 0)
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2] and outs[0][3] == outs[1][3]


@pytest.mark.parametrize("flags", [QOA_FUSE_FMA, QOA_FUSE_MERGE_DIAG, QOA_FUSE_FACTOR,
                                   QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG])
@pytest.mark.parametrize("tile", [6, 8, 12])
def test_fma_mode_within_tolerance(tile, flags):
    """FMA products (and merged diagonal runs, factored passes): not
    bit-identical, but every
    amplitude within 1e-12 of the oracle (original gate order) and fidelity
    >= 1 - 1e-12, on the layered circuit and on the full gate surface (2x2,
    4x4, controls)."""
    n = 16
    g = W.layered_random(n, 20, seed=2, ladders=True)
    gm, mats = mixed_circuit(n, 400, seed=tile)
    for gates, m in ((g, None), (gm, mats)):
        a0 = random_state(n, seed=tile)
        ref = a0.copy()
        po.run_gates(ref, gates, matrices=m)
        with StateVector(n) as sv:
            sv.upload(a0)
            sv.run_fused(gates, tile_qubits=tile, flags=flags, matrices=m)
            out = sv.download()
        assert np.max(np.abs(out - ref)) <= 1e-12
        assert fidelity(out, ref) >= 1 - 1e-12


@pytest.mark.parametrize("mode", MODES + [QOA_FUSE_FMA, QOA_FUSE_MERGE_DIAG,
                                   QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG,
                                   QOA_FUSE_RELABEL | QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG])
def test_c2_28q_fused_vs_per_gate_full_size(mode):
    """The full 28-qubit, 3164-gate circuit: fused passes against per-gate
    sweeps on the device (size-independent checks: unit norm, agreement
    <= 1e-12 in every amplitude, fidelity >= 1-1e-12)."""
    n = 28
    g = W.layered_random(n, 40, ladders=True)
    with StateVector(n) as a, StateVector(n) as b:
        a.run_gates(g)
        b.run_fused(g, flags=mode)
        na, nb = a.norm_sqr(), b.norm_sqr()
        assert abs(na - 1) < 1e-10 and abs(nb - 1) < 1e-10
        # compare in chunks on the host to bound memory
        da, db = a.download(), b.download()
    d = np.max(np.abs(da - db))
    f = fidelity(da, db)
    del da, db
    assert d <= 1e-12, d
    assert f >= 1 - 1e-12, f


def test_merged_diagonal_program_matches_emulator():
    """The merged program on the device against the CPU emulation of the same
    op program (the device multiplies with FMA, the emulator without: within
    1e-13), at 13 qubits with 10-qubit tiles (ladders give long runs)."""
    from tile_emulator import plan_pool, plan_program, run_program

    n, tile = 13, 10
    g = W.layered_random(n, 12, seed=9, ladders=True)
    ops, spos, rng, k = plan_program(n, g, tile_qubits=tile, flags=QOA_FUSE_MERGE_DIAG)
    pool = plan_pool(n, g, tile_qubits=tile, flags=QOA_FUSE_MERGE_DIAG)
    a0 = random_state(n, seed=3)
    emu = a0.copy()
    run_program(emu, n, ops, spos, rng, k, pool=pool)
    with StateVector(n) as sv:
        sv.upload(a0)
        sv.run_fused(g, tile_qubits=tile, flags=QOA_FUSE_MERGE_DIAG)
        out = sv.download()
    assert np.max(np.abs(out - emu)) <= 1e-13


@pytest.mark.parametrize("n,tile", [(16, 10), (17, 11), (18, 12)])
def test_relabelled_program_matches_oracle_exactly(n, tile):
    """Relabelling plans move qubits between positions from pass to pass and
    put them back at the end: exact mode stays bit-identical to the oracle in
    the planned order, and to the emulation of the same program."""
    from tile_emulator import plan_program, plan_store_positions, run_program

    g = W.layered_random(n, 16, seed=n + tile, ladders=True)
    ops, spos, rng, k = plan_program(n, g, tile_qubits=tile, flags=QOA_FUSE_RELABEL)
    st = plan_store_positions(n, g, tile_qubits=tile, flags=QOA_FUSE_RELABEL)
    assert np.any(st[:, :k] != spos[:, :k])
    a0 = random_state(n, seed=7)
    emu = a0.copy()
    run_program(emu, n, ops, spos, rng, k, spos_st=st)
    order, _ = plan_fused(n, g, tile_qubits=tile, flags=QOA_FUSE_RELABEL)
    ref = a0.copy()
    po.run_gates(ref, g, order=order)
    with StateVector(n) as sv:
        sv.upload(a0)
        sv.run_fused(g, tile_qubits=tile, flags=QOA_FUSE_RELABEL)
        out = sv.download()
    assert np.array_equal(out, ref) and np.array_equal(emu, ref)


def test_factored_program_matches_emulator():
    """The factored program (H as sums and differences, RX/RY with unit
    diagonals, RZ as half-register phases, one op per pass carrying the
    scalars) on the device against the CPU emulation of the same program."""
    from tile_emulator import plan_pool, plan_program, run_program

    n, tile = 14, 12
    flags = QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG
    g = W.layered_random(n, 12, seed=19, ladders=True)
    ops, spos, rng, k = plan_program(n, g, tile_qubits=tile, flags=flags)
    pool = plan_pool(n, g, tile_qubits=tile, flags=flags)
    a0 = random_state(n, seed=4)
    emu = a0.copy()
    run_program(emu, n, ops, spos, rng, k, pool=pool)
    with StateVector(n) as sv:
        sv.upload(a0)
        sv.run_fused(g, tile_qubits=tile, flags=flags)
        out = sv.download()
    assert np.max(np.abs(out - emu)) <= 1e-13
