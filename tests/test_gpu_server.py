"""Serving engine: continuous batching, admission, determinism alone vs mixed,
and the device sampler against its oracle mirror."""
import numpy as np
import pytest

import pyoracle as po
from kat_util import random_state
from qoa_amd import StateVector, plan_fused
from qoa_amd import workloads as W
from qoa_amd.server import Server, sample_state

pytestmark = pytest.mark.gpu

SAMPLE_KEY = 0x53414D504C45


@pytest.mark.parametrize("n", [3, 8, 12, 13, 17])
def test_sampler_matches_oracle(n):
    a = random_state(n, seed=n)
    with StateVector(n) as sv:
        sv.upload(a)
        got = sample_state(sv, seed=77, count=64)
    want = [po.lib().orc_sample(po._p(a), a.size, po.lib().orc_sample_uniform(77, 0, s))
            for s in range(64)]
    assert got.tolist() == want


def _job(seed, n):
    _, _, g = W.ansatz_job(seed, n_min=n, n_max=n, d_min=4, d_max=10)
    return g


def test_server_job_alone_equals_in_mix():
    """The same job (16 noisy trajectories) gives identical samples when it
    runs alone and when it shares every step with 30 other jobs, noisy and
    noise-free."""
    target_n, T = 11, 16
    g = _job(5, target_n)
    clean_n = 10  # a noise-free job: it runs inside the noisy launches in the mix
    gc = _job(6, clean_n)
    with Server(budget_bytes=2 << 30) as srv:
        jid = srv.submit(target_n, g, trajectories=T, p1=0.02, p2=0.05, noise_seed=99)
        alone = srv.wait(jid)
        assert alone.status == 0 and alone.samples.size == T
        clean_alone = srv.wait(srv.submit(clean_n, gc, trajectories=8, noise_seed=123))
        assert clean_alone.status == 0 and clean_alone.samples.size == 8
        ids = []
        for k in range(30):
            n = 6 + (k % 9)
            noisy = k % 3 != 0  # noise-free jobs share the noisy launches too
            ids.append(srv.submit(n, _job(100 + k, n), trajectories=8, p1=1e-3 if noisy else 0.0,
                                  p2=1e-2 if noisy else 0.0, noise_seed=k))
            if k == 12:
                jid2 = srv.submit(target_n, g, trajectories=T, p1=0.02, p2=0.05, noise_seed=99)
                jc2 = srv.submit(clean_n, gc, trajectories=8, noise_seed=123)
        mixed = srv.wait(jid2)
        clean_mixed = srv.wait(jc2)
        for i in ids:
            assert srv.wait(i).status == 0
        st = srv.stats()
    assert np.array_equal(alone.samples, mixed.samples)
    assert np.array_equal(clean_alone.samples, clean_mixed.samples)
    assert st["completed"] == 34 and st["fatal"] == 0


def test_server_samples_match_oracle_trajectories():
    """Each trajectory's sample equals the oracle trajectory (planned order,
    same Philox draws) followed by the fixed-order sampler."""
    n, T, seed = 9, 6, 4242
    g = _job(17, n)
    order, _ = plan_fused(n, g, tile_qubits=12)
    # the server plans in register mode with noise operands: same order as a
    # batch plan of this job; recompute through a one-job batch
    from qoa_amd.batch import Batch

    with Batch([(n, g)], tile_qubits=12, trajectories=T, p1=0.03, p2=0.08,
               noise_seed=seed) as b:
        order = b.order(0)
    with Server(budget_bytes=1 << 30) as srv:
        res = srv.wait(srv.submit(n, g, trajectories=T, p1=0.03, p2=0.08, noise_seed=seed))
    for t in range(T):
        a = po.state(n)
        po.run_trajectory(a, g, seed, t, 0.03, 0.08, order=order)
        u = po.lib().orc_sample_uniform(seed ^ SAMPLE_KEY, t, 0)
        assert res.samples[t] == po.lib().orc_sample(po._p(a), a.size, u), t


def test_server_admission_and_rejection():
    with Server(budget_bytes=64 << 20) as srv:  # 64 MiB arena
        big = srv.submit(24, _job(1, 24), trajectories=1)  # 256 MiB: can never fit
        r = srv.wait(big)
        assert r.status == -6
        ids = [srv.submit(18, _job(2 + k, 18), trajectories=4) for k in range(6)]  # 16 MiB each
        for i in ids:
            assert srv.wait(i, timeout=120).status == 0
        st = srv.stats()
        assert st["rejected"] == 1 and st["arena_peak"] <= 64 << 20
