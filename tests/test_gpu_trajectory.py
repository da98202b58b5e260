"""The plain per-gate trajectory loop (qoa_run_trajectory): every gate's
sweep followed by the drawn Pauli sweeps. Bit-identical to the oracle
trajectory in the original order, and, run in a batch plan's order, to the
batched in-kernel trajectory of the same job."""
import numpy as np
import pytest

import pyoracle as po
from qoa_amd import StateVector
from qoa_amd._native import QOA_ERR_ARG, GATE_DTYPE, OP, QoaError
from qoa_amd.batch import Batch
from test_gpu_batch import shared_table_jobs

pytestmark = pytest.mark.gpu


def test_plain_trajectory_matches_oracle_and_batch():
    p1, p2, T = 0.05, 0.2, 3
    jobs, table = shared_table_jobs([10, 13], 140, seed0=77)
    seeds = [501, 502]
    with Batch(jobs, tile_qubits=10, trajectories=T, matrices=table, p1=p1, p2=p2,
               noise_seed=seeds) as b:
        b.run()
        b.synchronize()
        for j, (n, g) in enumerate(jobs):
            order = b.order(j)
            with StateVector(n) as sv:
                for t in range(T):
                    sv.apply_reset_all()
                    sv.run_trajectory(g, seeds[j], t, p1, p2, matrices=table)
                    ref = po.state(n)
                    po.run_trajectory(ref, g, seeds[j], t, p1, p2, matrices=table)
                    assert np.array_equal(sv.download(), ref), (j, t, "original order")
                    sv.apply_reset_all()
                    sv.run_trajectory(g, seeds[j], t, p1, p2, matrices=table, order=order)
                    assert np.array_equal(sv.download(), b.download(j, t)), (j, t, "planned")


def test_plain_trajectory_rejects_bad_operands():
    g = np.array([(OP["CNOT"], 1, 1, 0, 0.0, 0)], dtype=GATE_DTYPE)
    with StateVector(4) as sv:
        with pytest.raises(QoaError) as e:
            sv.run_trajectory(g, 1, 0, 0.1, 0.1)
        assert e.value.code == QOA_ERR_ARG
        g2 = np.array([(OP["H"], 9, 0, 0, 0.0, 0)], dtype=GATE_DTYPE)
        with pytest.raises(QoaError):
            sv.run_trajectory(g2, 1, 0, 0.1, 0.1)
