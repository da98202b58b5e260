"""Full-size parity: the 28-qubit layered circuit with CPHASE ladders (3164
gates, the headline workload) and batched ansatz jobs of 20-24 qubits and
depth >= 40, each against the oracle run on all host cores (the gate loops
split over threads are the same arithmetic, tests/test_oracle_invariants.py).

  per-gate sweeps  bit-identical to the oracle (src/vectorization.rs:904-1261)
  fused passes     max |d amp| <= 1e-12 and 1 - F <= 1e-12 against the
                   oracle in the circuit's own order (exact, FMA and merged
                   diagonal modes)
  batched jobs     bit-identical to the oracle in each job's planned order
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np
import pytest

import pyoracle as po
from qoa_amd import StateVector
from qoa_amd import workloads as W
from qoa_amd._native import (QOA_FUSE_FACTOR, QOA_FUSE_FMA, QOA_FUSE_MERGE_DIAG, QOA_FUSE_RELABEL,
                             QOA_FUSE_SPECIALIZE)
from qoa_amd.batch import Batch

pytestmark = pytest.mark.gpu

THREADS = max(1, min(os.cpu_count() or 1, 16))


def say(msg: str) -> None:
    print(f"[fullsize] {msg}", file=sys.stderr, flush=True)


def oracle_run(n, g, order=None):
    ref = po.state(n)
    po.set_threads(THREADS)
    try:
        po.run_gates(ref, g, order=order)
    finally:
        po.set_threads(1)
    return ref


def errors(a: np.ndarray, ref: np.ndarray) -> tuple[float, float]:
    d = float(np.max(np.abs(a - ref)))
    f = abs(np.vdot(ref, a)) ** 2 / (np.vdot(ref, ref).real * np.vdot(a, a).real)
    return d, 1.0 - f


def test_c2_28_qubits_against_oracle():
    n = 28
    g = W.layered_random(n, 40, seed=20241203, ladders=True)
    assert len(g) == 3164
    t0 = time.perf_counter()
    say(f"oracle: 28 qubits, {len(g)} gates, {THREADS} threads")
    ref = oracle_run(n, g)
    say(f"oracle done in {time.perf_counter() - t0:.1f} s")
    with StateVector(n) as sv:
        sv.run_gates(g)
        out = sv.download()
        assert np.array_equal(out, ref), "per-gate sweeps differ from the oracle"
        del out
        say("per-gate: bit-identical")
        merged = None
        for name, flags in (("exact", 0), ("fma", QOA_FUSE_FMA),
                            ("merged diagonals", QOA_FUSE_MERGE_DIAG),
                            ("merged diagonals, specialised passes",
                             QOA_FUSE_MERGE_DIAG | QOA_FUSE_SPECIALIZE),
                            ("factored, merged diagonals", QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG),
                            ("factored, merged diagonals, specialised passes",
                             QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG | QOA_FUSE_SPECIALIZE),
                            ("relabelled, factored, merged diagonals",
                             QOA_FUSE_RELABEL | QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG),
                            ("relabelled, factored, merged diagonals, specialised passes",
                             QOA_FUSE_RELABEL | QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG
                             | QOA_FUSE_SPECIALIZE)):
            sv.apply_reset_all()
            t1 = time.perf_counter()
            sv.run_fused(g, flags=flags)
            sv.synchronize()
            out = sv.download()
            d, inf = errors(out, ref)
            say(f"fused {name}: max|d amp| {d:.3e}, 1-F {inf:.3e} "
                f"({time.perf_counter() - t1:.1f} s with planning and compiling)")
            assert d <= 1e-12 and inf <= 1e-12, (name, d, inf)
            if not flags & QOA_FUSE_SPECIALIZE:
                merged = out
            elif flags & QOA_FUSE_MERGE_DIAG:  # the op-loop kernel's bits exactly
                assert np.array_equal(out, merged)
            del out


def test_large_ansatz_jobs_batched_against_oracle():
    jobs = []
    for seed in (11, 12, 13, 14):
        n, depth, g = W.ansatz_job(9000 + seed, n_min=20, n_max=24, d_min=40, d_max=60)
        jobs.append((n, g))
    say("batch jobs (n, gates): " + ", ".join(f"({n}, {len(g)})" for n, g in jobs))
    with Batch(jobs, tile_qubits=12) as b:
        b.run()
        b.synchronize()
        for j, (n, g) in enumerate(jobs):
            ref = oracle_run(n, g, order=b.order(j))
            assert np.array_equal(b.download(j), ref), (j, n)
            d, inf = errors(b.download(j), oracle_run(n, g))
            say(f"job {j}: {n} qubits bit-identical in planned order; original order "
                f"max|d| {d:.2e}, 1-F {inf:.2e}")
            assert d <= 1e-12 and inf <= 1e-12
