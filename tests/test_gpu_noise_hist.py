"""Noisy trajectories on the GPU against the exact noisy distribution.

The serving engine runs many trajectories of one circuit with the in-kernel
Pauli channel (p1 after one-qubit gates, p2 on every operand of larger gates)
and returns one Z-basis sample per trajectory. Their histogram must match the
distribution a numpy density-matrix run gives (tests/noise_dm.py: textbook
gate matrices, the channel as a superoperator) within 5 sigma of the
multinomial TVD bound. Each test also shows, on the CPU, that deliberately
wrong channels lie outside that bound, so the comparison has the power to see
them. The oracle is not involved: this pins the device channel on its own.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import noise_dm as dm
from qoa_amd.batch import Batch
from qoa_amd.server import Server
from test_oracle_trajectory import ensemble_circuit, gates

pytestmark = pytest.mark.gpu

WRONG = ("no_noise", "first_operand_only", "no_y", "p_swapped")


def circuit10() -> np.ndarray:
    """Two ensemble layers on 10 qubits plus gates between far-apart qubits,
    so that with 8-qubit tiles errors land on qubits outside a pass's tile."""
    extra = gates([("CNOT", (9, 3, 0), 0.0), ("H", (8, 0, 0), 0.0), ("CZ", (2, 7, 0), 0.0),
                   ("TOFFOLI", (9, 5, 0), 0.0), ("RX", (6, 0, 0), 0.8), ("CNOT", (4, 8, 0), 0.0)])
    return np.concatenate([ensemble_circuit(10, 2), extra])


def server_samples(n, g, p1, p2, jobs, traj, tile, seed0):
    out = []
    with Server(budget_bytes=4 << 30, tile_qubits=tile) as srv:
        ids = [srv.submit(n, g, trajectories=traj, p1=p1, p2=p2, noise_seed=seed0 + j)
               for j in range(jobs)]
        for i in ids:
            r = srv.wait(i, timeout=120)
            assert r.status == 0 and r.samples.size == traj
            out.append(r.samples)
        assert srv.stats()["fatal"] == 0
    return np.concatenate(out)


def test_server_histogram_4_qubits():
    n, p1, p2 = 4, 0.01, 0.1
    g = ensemble_circuit(n, 2)
    exact = dm.noisy_probabilities(n, g, p1, p2)
    s = server_samples(n, g, p1, p2, jobs=64, traj=1024, tile=12, seed0=5000)
    hist = np.bincount(s.astype(np.int64), minlength=1 << n) / s.size
    bound = dm.tvd_bound(exact, s.size)
    print(f"4q: TVD {dm.tvd(hist, exact):.4f} bound {bound:.4f} samples {s.size}", flush=True)
    assert dm.tvd(hist, exact) < bound
    for w in WRONG:
        assert dm.tvd(dm.noisy_probabilities(n, g, p1, p2, w), exact) > 2 * bound, w


def test_server_histogram_10_qubits_multi_pass():
    """8-qubit tiles: several passes per trajectory, errors on qubits
    outside the tile go through the X frame. Marginals over two 4-qubit
    sets (inside and across tiles) are compared."""
    n, p1, p2 = 10, 0.01, 0.1
    g = circuit10()
    exact = dm.noisy_probabilities(n, g, p1, p2)
    wrong = {w: dm.noisy_probabilities(n, g, p1, p2, w) for w in WRONG}
    s = server_samples(n, g, p1, p2, jobs=64, traj=1024, tile=8, seed0=9000)
    for sub in ([6, 7, 8, 9], [3, 5, 7, 9]):
        me = dm.marginal(exact, n, sub)
        got = dm.marginal_of_samples(s, sub)
        bound = dm.tvd_bound(me, s.size)
        print(f"10q {sub}: TVD {dm.tvd(got, me):.4f} bound {bound:.4f}", flush=True)
        assert dm.tvd(got, me) < bound, sub
        for w, pw in wrong.items():
            assert dm.tvd(dm.marginal(pw, n, sub), me) > 2 * bound, (w, sub)


def test_batch_mean_distribution_10_qubits():
    """Batched noisy trajectories (8-qubit tiles): the mean of the
    trajectories' outcome distributions equals the exact one within 5 sigma
    per outcome."""
    n, p1, p2, T = 10, 0.01, 0.1, 1024
    g = circuit10()
    exact = dm.noisy_probabilities(n, g, p1, p2)
    with Batch([(n, g)], tile_qubits=8, trajectories=T, p1=p1, p2=p2, noise_seed=77) as b:
        b.run()
        b.synchronize()
        probs = np.array([np.abs(b.download(0, t)) ** 2 for t in range(T)])
    mean = probs.mean(axis=0)
    sigma = probs.std(axis=0) / math.sqrt(T) + 1e-9
    z = np.abs(mean - exact) / sigma
    print(f"batch 10q: max z {z.max():.2f}", flush=True)
    assert z.max() < 5.0
    zw = np.abs(mean - dm.noisy_probabilities(n, g, p1, p2, "no_noise")) / sigma
    assert zw.max() > 10.0
