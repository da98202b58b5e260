"""The fused-pass planner (host code, no GPU): its executed order is a valid
reordering of the circuit, checked by running the oracle in both orders."""
import numpy as np
import pytest

import pyoracle as po
from kat_util import random_state
from qoa_amd import _native as N
from qoa_amd import plan_fused
from qoa_amd import workloads as W
from qoa_amd._native import GATE_DTYPE, OP, QOA_CZ_REF_MASKS, QOA_FUSE_FACTOR, QOA_FUSE_MERGE_DIAG


def mixed_circuit(n: int, count: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Every gate kind of the surface, random operands (some out of range)."""
    rng = np.random.default_rng(seed)
    mats = []
    rows = []
    kinds = ["H", "X", "Y", "Z", "S", "T", "PHASE", "RX", "RY", "RZ", "CNOT", "CZ", "CPHASE",
             "SWAP", "CSWAP", "TOFFOLI", "CY", "U2", "U4"]
    for _ in range(count):
        k = kinds[rng.integers(len(kinds))]
        qs = rng.choice(n, size=3, replace=False)
        q0, q1, q2 = (int(x) for x in qs)
        th = float(rng.uniform(0, 2 * np.pi))
        aux = 0
        if k in ("H", "RZ", "X") and rng.random() < 0.05:
            q0 = n + 1  # out of range: no-op (RZ: global phase)
        if k == "U2":
            u = np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))[0]
            aux = len(mats) // 1
            m = np.zeros(16, dtype=np.complex128)
            m[:4] = u.reshape(4)
            mats.append(m)
        if k == "U4":
            u = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]
            aux = len(mats)
            mats.append(u.reshape(16))
        rows.append((OP[k], q0, q1, q2, th, aux))
    g = np.array(rows, dtype=GATE_DTYPE)
    # U2 reads 8 doubles at 8*aux and U4 32 doubles at 32*aux: lay matrices out
    # so both indexings hit the right entries (one 16-complex slot per matrix).
    table = np.zeros(16 * max(1, len(mats)), dtype=np.complex128)
    for i, m in enumerate(mats):
        table[16 * i:16 * i + 16] = m
    for i in range(len(g)):
        if g[i]["op"] == OP["U2"]:
            g[i]["aux"] = 4 * int(g[i]["aux"])
    return g, table


@pytest.mark.parametrize("flags", [0, 2])
@pytest.mark.parametrize("tile", [4, 6, 8, 12])
def test_plan_is_permutation_and_valid_reordering(tile, flags):
    n = 10
    g, mats = mixed_circuit(n, 300, seed=tile)
    order, passes = plan_fused(n, g, tile_qubits=tile, flags=flags)
    assert sorted(order.tolist()) == list(range(len(g)))
    assert passes >= 1
    a0 = random_state(n, 3)
    ref = a0.copy()
    po.run_gates(ref, g, matrices=mats)
    alt = a0.copy()
    po.run_gates(alt, g, matrices=mats, order=order)
    assert np.max(np.abs(alt - ref)) < 1e-12


@pytest.mark.parametrize("flags", [0, QOA_CZ_REF_MASKS])
def test_plan_c2_shape(flags):
    n = 14
    g = W.layered_random(n, 15, seed=5, ladders=True)
    order, passes = plan_fused(n, g, tile_qubits=8, flags=flags)
    ref = po.state(n)
    po.run_gates(ref, g, flags=flags)
    alt = po.state(n)
    po.run_gates(alt, g, flags=flags, order=order)
    assert np.max(np.abs(alt - ref)) < 1e-12


def test_plan_whole_state_in_one_tile():
    g = W.layered_random(9, 20, seed=1, ladders=True)
    order, passes = plan_fused(9, g, tile_qubits=12, flags=2)  # LDS-resident: no reordering
    assert passes == 1 and order.tolist() == list(range(len(g)))
    order, passes = plan_fused(9, g, tile_qubits=12)  # register segments reorder in the tile
    assert passes == 1 and sorted(order.tolist()) == list(range(len(g)))
    ref = po.state(9)
    po.run_gates(ref, g)
    alt = po.state(9)
    po.run_gates(alt, g, order=order)
    assert np.max(np.abs(alt - ref)) < 1e-12


def test_plan_pass_counts_28q():
    """The 3164-gate 28-qubit circuit needs far fewer passes than gates."""
    g = W.layered_random(28, 40, ladders=True)
    _, passes = plan_fused(28, g, tile_qubits=12, low_qubits=3)
    assert passes < len(g) / 8, passes


# ---- emulated op programs (CPU): the planner's output is checked as the
# device kernels execute it, against the oracle run in the planned order ----
@pytest.mark.parametrize("flags", [0, QOA_CZ_REF_MASKS, 2, 2 | QOA_CZ_REF_MASKS])
@pytest.mark.parametrize("n,tile", [(9, 0), (12, 8), (11, 10)])
def test_emulated_program_matches_oracle(n, tile, flags):
    from tile_emulator import plan_program, run_program

    g, mats = mixed_circuit(n, 250, seed=n * 3 + tile)
    g = g[g["op"] != OP["U4"]]
    a0 = random_state(n, seed=n + tile)
    ops, spos, rng, k = plan_program(n, g, tile_qubits=tile, flags=flags, matrices=mats)
    order, _ = plan_fused(n, g, tile_qubits=tile, flags=flags)
    ref = a0.copy()
    po.run_gates(ref, g, flags=flags, matrices=mats, order=order)
    a = a0.copy()
    run_program(a, n, ops, spos, rng, k)
    assert np.array_equal(a, ref)


@pytest.mark.parametrize("tile", [8, 12])
def test_emulated_ansatz_and_layers(tile):
    from tile_emulator import plan_program, run_program

    for n, g in ((12, W.ansatz_job(900, n_min=12, n_max=12, d_min=2, d_max=12)[2]),
                 (13, W.layered_random(13, 10, seed=4, ladders=True))):
        ops, spos, rng, k = plan_program(n, g, tile_qubits=tile)
        order, _ = plan_fused(n, g, tile_qubits=tile)
        ref = po.state(n)
        po.run_gates(ref, g, order=order)
        a = po.state(n)
        run_program(a, n, ops, spos, rng, k)
        assert np.array_equal(a, ref)


@pytest.mark.parametrize("n,tile,ladders", [(13, 10, True), (14, 12, True), (12, 8, False)])
def test_merged_diagonal_runs_match_oracle(n, tile, ladders):
    """QOA_FUSE_MERGE_DIAG: runs of diagonal ops become one op whose factors
    are multiplied per register group first. The emulated program equals the
    oracle in the planned order within 1e-12 (products regrouped, not
    bit-identical), merged ops exist, and they carry every diagonal gate."""
    from tile_emulator import TK_DIAGN, plan_pool, plan_program, run_program

    if ladders:
        g = W.layered_random(n, 10, seed=n, ladders=True)
        mats = None
    else:
        g, mats = mixed_circuit(n, 300, seed=n)
        g = g[g["op"] != OP["U4"]]
    flags = QOA_FUSE_MERGE_DIAG
    ops, spos, rng, k = plan_program(n, g, tile_qubits=tile, flags=flags, matrices=mats)
    pool = plan_pool(n, g, tile_qubits=tile, flags=flags, matrices=mats)
    plain, _, _, _ = plan_program(n, g, tile_qubits=tile, matrices=mats)
    merged = ops[ops["kind"] == TK_DIAGN]
    assert len(merged) > 0 and len(ops) < len(plain)
    order, _ = plan_fused(n, g, tile_qubits=tile, flags=flags)
    ref = random_state(n, seed=n)
    a = ref.copy()
    po.run_gates(ref, g, matrices=mats, order=order)
    run_program(a, n, ops, spos, rng, k, pool=pool)
    assert np.max(np.abs(a - ref)) <= 1e-12
    orig = random_state(n, seed=n)
    po.run_gates(orig, g, matrices=mats)
    assert np.max(np.abs(a - orig)) <= 1e-12


@pytest.mark.parametrize("n,tile,extra", [(13, 10, QOA_FUSE_MERGE_DIAG), (14, 12, 0), (12, 8, 0)])
def test_factored_passes_match_oracle(n, tile, extra):
    """QOA_FUSE_FACTOR: uncontrolled H, RX/RY and register-slot RZ lose their
    scalar factor (H sums and differences, RX/RY with m00 = m11 = 1, RZ a
    phase on the registers of its slot), and one op per pass takes the
    product. The emulated program equals the oracle in the planned and the
    original order within 1e-12; a lost or doubled factor (1/sqrt 2 of an H,
    cos of a rotation, e^{-i theta/2} of an RZ) would not."""
    from tile_emulator import TK_DPHASE, TK_H, TK_M, TK_RZ, plan_pool, plan_program, run_program

    g = W.layered_random(n, 10, seed=n + 7, ladders=True)
    flags = QOA_FUSE_FACTOR | extra
    ops, spos, rng, k = plan_program(n, g, tile_qubits=tile, flags=flags)
    pool = plan_pool(n, g, tile_qubits=tile, flags=flags)
    plain, _, _, _ = plan_program(n, g, tile_qubits=tile, flags=extra)
    h_raw = (ops["kind"] == TK_H) & (ops["c"][:, 0] == 1.0) & (ops["c"][:, 1] == 0.0)
    u = (ops["kind"] == TK_M) & np.isin(ops["t2"], (3, 4))
    assert h_raw.sum() > 0 and u.sum() > 0
    assert np.all(ops[u]["c"][:, 0] == 1.0) and np.all(ops[u]["c"][:, 6] == 1.0)
    if not extra:  # RZ on register slots became phases on half the registers
        assert (ops["kind"] == TK_RZ).sum() < (plain["kind"] == TK_RZ).sum()
        assert (ops["kind"] == TK_DPHASE).sum() > (plain["kind"] == TK_DPHASE).sum()
    order, _ = plan_fused(n, g, tile_qubits=tile, flags=flags)
    ref = random_state(n, seed=n)
    a = ref.copy()
    po.run_gates(ref, g, order=order)
    run_program(a, n, ops, spos, rng, k, pool=pool)
    assert np.max(np.abs(a - ref)) <= 1e-12
    orig = random_state(n, seed=n)
    po.run_gates(orig, g)
    assert np.max(np.abs(a - orig)) <= 1e-12


def test_factored_growth_stays_bounded():
    """Rotations near pi have tiny cosines: taking their factor out multiplies
    the amplitudes by 1/|cos| inside the pass. The planner stops taking
    factors before the product passes 2^64, so forty RY(pi - 1e-9) on one
    qubit stay finite and within 1e-12 of the oracle."""
    from tile_emulator import TK_M, plan_pool, plan_program, run_program

    n = 10
    rows = [(OP["H"], q, 0, 0, 0.0, 0) for q in range(n)]
    rows += [(OP["RY"], 3, 0, 0, np.pi - 1e-9, 0)] * 40
    rows += [(OP["RX"], 4, 0, 0, np.pi, 0)] * 10
    g = np.array(rows, dtype=GATE_DTYPE)
    ops, spos, rng, k = plan_program(n, g, tile_qubits=8, flags=QOA_FUSE_FACTOR)
    factored = (ops["kind"] == TK_M) & np.isin(ops["t2"], (3, 4))
    assert 0 < factored.sum() < 50
    a = random_state(n, seed=5)
    ref = a.copy()
    run_program(a, n, ops, spos, rng, k, pool=plan_pool(n, g, tile_qubits=8, flags=QOA_FUSE_FACTOR))
    po.run_gates(ref, g)
    assert np.all(np.isfinite(a)) and np.max(np.abs(a - ref)) <= 1e-12


@pytest.mark.parametrize("n,tile,kind", [(14, 10, "layered"), (15, 12, "layered"),
                                         (13, 10, "mixed"), (14, 12, "mixed"), (16, 12, "ladder")])
def test_relabelled_plans_match_oracle(n, tile, kind):
    """QOA_FUSE_RELABEL: passes store their tile with qubits at other tile
    positions, so the next pass finds its own qubits at the three low
    positions. Pure data movement: the emulated program (loads and stores at
    the plan's positions) equals the oracle in the planned order bit for bit,
    the state ends in the original qubit order, every register pass loads and
    stores the three low positions from lanes 0..2, and passes do relabel."""
    from tile_emulator import plan_program, plan_store_positions, run_program

    mats = None
    if kind == "mixed":  # U4 (MAT4) runs on the GPU only
        g, mats = mixed_circuit(n, 400, seed=n + tile)
        g = g[g["op"] != OP["U4"]]
    else:
        g = W.layered_random(n, 16, seed=n, ladders=kind == "ladder")
    flags = N.QOA_FUSE_RELABEL
    ops, spos, rng, k = plan_program(n, g, tile_qubits=tile, flags=flags, matrices=mats)
    st = plan_store_positions(n, g, tile_qubits=tile, flags=flags, matrices=mats)
    assert st.shape == spos.shape
    assert np.any(st[:, :k] != spos[:, :k]), "no pass relabels"
    order, _ = plan_fused(n, g, tile_qubits=tile, flags=flags)
    ref = random_state(n, seed=n + 1)
    a = ref.copy()
    po.run_gates(ref, g, matrices=mats, order=order)
    run_program(a, n, ops, spos, rng, k, spos_st=st)
    assert np.array_equal(a, ref)


def test_relabelled_factored_merged_plan_matches_oracle():
    from tile_emulator import plan_pool, plan_program, plan_store_positions, run_program

    n, tile = 15, 12
    g = W.layered_random(n, 20, seed=31, ladders=True)
    flags = N.QOA_FUSE_RELABEL | QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG
    ops, spos, rng, k = plan_program(n, g, tile_qubits=tile, flags=flags)
    st = plan_store_positions(n, g, tile_qubits=tile, flags=flags)
    pool = plan_pool(n, g, tile_qubits=tile, flags=flags)
    a = random_state(n, seed=2)
    ref = a.copy()
    run_program(a, n, ops, spos, rng, k, pool=pool, spos_st=st)
    po.run_gates(ref, g)
    assert np.max(np.abs(a - ref)) <= 1e-12


def test_relabelling_cuts_passes_of_the_28_qubit_circuit():
    """The 28-qubit layered circuit (3164 gates): fewer passes when the low
    positions follow the next pass's qubits (35 with qubits 0..2 in every
    tile)."""
    from tile_emulator import plan_program

    g = W.layered_random(28, 40, seed=20241203, ladders=True)
    base = plan_program(28, g, flags=QOA_FUSE_MERGE_DIAG)[2]
    rel = plan_program(28, g, flags=QOA_FUSE_MERGE_DIAG | N.QOA_FUSE_RELABEL)[2]
    assert len(rel) < len(base)


def test_sincos_coefficients_match_oracle():
    """Gate coefficients come from sincos on both sides (it can differ from
    separate cos/sin in the last bit, e.g. at theta = 3.3790254426509847)."""
    from tile_emulator import plan_program

    th = 3.3790254426509847
    g = np.array([(OP["RY"], 0, 0, 0, th, 0)], dtype=GATE_DTYPE)
    ops, _, _, _ = plan_program(1, g, tile_qubits=4)
    a = np.array([1, 0], dtype=complex)
    po.orc.ry(a, 0, th)
    assert ops[0]["c"][0] == a[0].real and ops[0]["c"][4] == a[1].real
