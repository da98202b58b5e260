"""Batched circuits (one launch per pass round): every job bit-identical to
the oracle in its planned order, within 1e-12 of its original order."""
import numpy as np
import pytest

import pyoracle as po
from qoa_amd import _native
from qoa_amd import workloads as W
from qoa_amd.batch import Batch
from test_planner import mixed_circuit

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tile", [8, 10, 12])
def test_ansatz_batch_matches_oracle(tile):
    jobs = []
    for j in range(24):
        n, depth, g = W.ansatz_job(900 + j, n_min=4, n_max=15, d_min=2, d_max=12)
        jobs.append((n, g))
    with Batch(jobs, tile_qubits=tile) as b:
        b.run()
        ms = b.synchronize()
        assert ms > 0
        info = b.info()
        assert info["rounds"] >= 1
        first = [b.download(j) for j in range(len(jobs))]
        for j, (n, g) in enumerate(jobs):
            out = first[j]
            ref = po.state(n)
            po.run_gates(ref, g, order=b.order(j))
            assert np.array_equal(out, ref), f"job {j} n={n}"
            orig = po.state(n)
            po.run_gates(orig, g)
            assert np.max(np.abs(out - orig)) <= 1e-12
        # a second run starts again from |0...0>
        b.run()
        b.synchronize()
        for j in range(len(jobs)):
            assert np.array_equal(b.download(j), first[j])


@pytest.mark.parametrize("merge", [False, True])
@pytest.mark.parametrize("tile", [8, 10, 12])
def test_factored_ansatz_batch_within_tolerance(tile, merge):
    """QOA_FUSE_FACTOR for batches: RY as [[1, -t], [t, 1]] with the cosine
    applied once per pass, RZ as a phase on half the registers; every job
    within 1e-12 of the oracle (original order); with QOA_FUSE_MERGE_DIAG
    too (runs of RZ phases and CZ signs merged, FMA kernels)."""
    jobs = []
    for j in range(24):
        n, depth, g = W.ansatz_job(700 + j, n_min=4, n_max=15, d_min=2, d_max=12)
        jobs.append((n, g))
    flags = _native.QOA_FUSE_FACTOR | (_native.QOA_FUSE_MERGE_DIAG if merge else 0)
    with Batch(jobs, tile_qubits=tile, flags=flags) as b:
        b.run()
        b.synchronize()
        for j, (n, g) in enumerate(jobs):
            out = b.download(j)
            orig = po.state(n)
            po.run_gates(orig, g)
            assert np.max(np.abs(out - orig)) <= 1e-12, f"job {j} n={n}"


def test_mixed_surface_batch_and_trajectory_copies():
    jobs = []
    mats_all = []
    for j, n in enumerate([5, 9, 13, 14]):
        g, mats = mixed_circuit(n, 150, seed=j)
        g = g[(g["op"] != 0xE2) & (g["op"] != 0xE3)]  # matrices differ per job: drop U2/U4
        jobs.append((n, g))
    with Batch(jobs, tile_qubits=10, trajectories=3) as b:
        b.run()
        b.synchronize()
        for j, (n, g) in enumerate(jobs):
            ref = po.state(n)
            po.run_gates(ref, g, order=b.order(j))
            for t in range(3):
                assert np.array_equal(b.download(j, t), ref), (j, t)


@pytest.mark.parametrize("tile", [8, 10, 12])
def test_noisy_trajectories_match_oracle(tile):
    """In-kernel Philox noise: each trajectory equals the oracle trajectory
    (same draws, planned order) bit for bit, including errors on qubits
    outside a pass's tile (X frame) for the multi-pass jobs."""
    p1, p2 = 0.05, 0.2
    jobs = []
    for j, n in enumerate([5, 9, 12, 14, 16]):
        _, _, g = W.ansatz_job(300 + j, n_min=n, n_max=n, d_min=6, d_max=6)
        jobs.append((n, g))
    seeds = [1000 + j for j in range(len(jobs))]
    T = 4
    with Batch(jobs, tile_qubits=tile, trajectories=T, p1=p1, p2=p2, noise_seed=seeds) as b:
        b.run()
        b.synchronize()
        for j, (n, g) in enumerate(jobs):
            order = b.order(j)
            outs = []
            for t in range(T):
                ref = po.state(n)
                po.run_trajectory(ref, g, seeds[j], t, p1, p2, order=order)
                out = b.download(j, t)
                assert np.array_equal(out, ref), (j, n, t)
                outs.append(out)
            assert any(not np.array_equal(outs[0], o) for o in outs[1:]), "noise had no effect"


@pytest.mark.parametrize("tile", [8, 12])
def test_factored_noisy_trajectories_within_tolerance(tile):
    """QOA_FUSE_FACTOR with trajectory noise: the pass scalars commute with the
    Pauli errors, so every factored trajectory stays within 1e-12 of the
    oracle trajectory with the same draws (original order)."""
    p1, p2 = 0.05, 0.2
    jobs = []
    for j, n in enumerate([5, 9, 12, 14, 16]):
        _, _, g = W.ansatz_job(500 + j, n_min=n, n_max=n, d_min=6, d_max=6)
        jobs.append((n, g))
    seeds = [2000 + j for j in range(len(jobs))]
    T = 3
    with Batch(jobs, tile_qubits=tile, trajectories=T, p1=p1, p2=p2, noise_seed=seeds,
               flags=_native.QOA_FUSE_FACTOR) as b:
        b.run()
        b.synchronize()
        for j, (n, g) in enumerate(jobs):
            order = b.order(j)
            for t in range(T):
                ref = po.state(n)
                po.run_trajectory(ref, g, seeds[j], t, p1, p2, order=order)
                assert np.max(np.abs(b.download(j, t) - ref)) <= 1e-12, (j, n, t)


def shared_table_jobs(sizes, count, seed0):
    """mixed_circuit jobs (every gate kind, U2/U4 included) with in-range
    operands, their matrices moved into one shared table."""
    jobs, tables, slot = [], [], 0
    for j, n in enumerate(sizes):
        g, mats = mixed_circuit(n, count, seed=seed0 + j)
        g = g[g["q0"] < n].copy()  # trajectories need in-range operands
        nslots = mats.size // 16
        u2 = g["op"] == 0xE2
        u4 = g["op"] == 0xE3
        g["aux"][u2] += 4 * slot
        g["aux"][u4] += slot
        tables.append(mats)
        slot += nslots
        jobs.append((n, g))
    return jobs, np.ascontiguousarray(np.concatenate(tables))


@pytest.mark.parametrize("tile", [8, 10])
def test_noisy_mixed_surface_matches_oracle(tile):
    """Noisy trajectories of circuits with every gate kind (Toffoli/CSWAP
    with three operands, X-type ops with controls outside the tile under the
    X frame, H, CY, U2/U4): each trajectory equals the oracle trajectory in
    planned order bit for bit, for jobs that take several passes."""
    p1, p2, T = 0.05, 0.2, 3
    jobs, table = shared_table_jobs([12, 14, 16], 160, seed0=40 + tile)
    seeds = [7000 + j for j in range(len(jobs))]
    with Batch(jobs, tile_qubits=tile, trajectories=T, matrices=table, p1=p1, p2=p2,
               noise_seed=seeds) as b:
        b.run()
        b.synchronize()
        for j, (n, g) in enumerate(jobs):
            order = b.order(j)
            for t in range(T):
                ref = po.state(n)
                po.run_trajectory(ref, g, seeds[j], t, p1, p2, matrices=table, order=order)
                assert np.array_equal(b.download(j, t), ref), (tile, j, n, t)


def test_graph_replays_equal_direct_launches():
    """The batch's launch sequence captured as a graph: several replays, with
    downloads and synchronisations between them, equal the run launched
    kernel by kernel (and the oracle in planned order)."""
    from qoa_amd._native import load

    lib = load()
    jobs = []
    for j in range(12):
        n, depth, g = W.ansatz_job(1300 + j, n_min=6, n_max=15, d_min=3, d_max=10)
        jobs.append((n, g))
    outs = {}
    for graphs in (0, 1):
        assert lib.qoa_set_option(b"batch_graphs", graphs) == 0
        try:
            with Batch(jobs, tile_qubits=10) as b:
                runs = []
                for _ in range(3):
                    b.run()
                    b.synchronize()
                    runs.append([b.download(j) for j in range(len(jobs))])
                for r in runs[1:]:
                    for j in range(len(jobs)):
                        assert np.array_equal(r[j], runs[0][j]), (graphs, j)
                outs[graphs] = runs[0]
                for j, (n, g) in enumerate(jobs):
                    ref = po.state(n)
                    po.run_gates(ref, g, order=b.order(j))
                    assert np.array_equal(runs[0][j], ref), (graphs, j)
        finally:
            lib.qoa_set_option(b"batch_graphs", 1)
    for j in range(len(jobs)):
        assert np.array_equal(outs[0][j], outs[1][j])


def test_autotuned_tiles_are_valid_and_exact():
    from qoa_amd._native import QOA_BATCH_AUTOTUNE

    jobs = []
    for j in range(10):
        n, depth, g = W.ansatz_job(1400 + j, n_min=9, n_max=16, d_min=4, d_max=12)
        jobs.append((n, g))
    with Batch(jobs, tile_qubits=12, flags=QOA_BATCH_AUTOTUNE) as b:
        b.run()
        b.synchronize()
        for j, (n, g) in enumerate(jobs):
            tile, passes = b.job_plan(j)
            assert (tile == n if n <= 10 else 10 <= tile <= min(12, n)) and passes >= 1
            ref = po.state(n)
            po.run_gates(ref, g, order=b.order(j))
            assert np.array_equal(b.download(j), ref), (j, n, tile)


def test_fp32_block_batch_within_tolerance():
    """The FP32-fused batch (5-qubit blocks on matrix cores, complex64 states):
    every job within 1 - F <= 1e-5 and max |d amp| <= 1e-5 of the oracle's
    FP64 run; a second run starts again from |0...0>."""
    from qoa_amd._native import QOA_C64

    jobs = []
    for j in range(20):
        n, depth, g = W.ansatz_job(1500 + j, n_min=5, n_max=16, d_min=3, d_max=20)
        jobs.append((n, g))
    with Batch(jobs, precision=QOA_C64) as b:
        for rep in range(2):
            b.run()
            b.synchronize()
            for j, (n, g) in enumerate(jobs):
                out = b.download(j)
                ref = po.state(n)
                po.run_gates(ref, g)
                f = abs(np.vdot(ref, out)) ** 2 / (np.vdot(ref, ref).real * np.vdot(out, out).real)
                assert 1 - f <= 1e-5 and np.max(np.abs(out - ref)) <= 1e-5, (rep, j, n, 1 - f)
                assert b.job_plan(j)[0] == 5 and b.job_plan(j)[1] >= 1
