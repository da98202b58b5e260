"""QOA assembly programs on the device against the same interpreter driving
the oracle's QuantumState restatement (same Philox seed): equal measurement
outcomes and output, amplitudes within 1e-12."""
import numpy as np
import pytest

from qoa_amd import program as P
from qoa_amd.state import NoiseConfig
from test_program import OracleBackend

pytestmark = pytest.mark.gpu

PROGRAM = """
; teleport-style chain on 7 qubits, every instruction kind of the subset
QINIT 6
RZ 0 3
H 0
RZ 0 6
HAD 1
CZ 1 2
CZ 0 1
APPLYHADAMARD 0
QMEAS 0
QMEAS 1
APPLYBITFLIP 2
RX 3 0.7
RY 4 1.3
CNOT 3 4
CPHASE 4 5 0.785398
SWAP 5 6
CSWAP 3 5 6
T 6
S 5
P 4 0.25
PHASE 4 9.0          ; read and ignored by the VM
RESET 3              ; read and ignored by the VM
Z 2
CHARLOAD 2 'Q'
RSA                  ; swallowed by CHARLOAD's 4-byte advance in the VM
CHAROUT 5
MEAS 6
VERBOSELOG 0 "done"
"""


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_program_matches_oracle_backend(seed):
    qs, res = P.run_text(PROGRAM, seed=seed)
    with qs:
        out = qs.download()
    ob = OracleBackend(res.n_qubits, seed)
    ref = P.execute(P.compile_text(PROGRAM).payload, ob)
    assert res.n_qubits == 7
    assert res.measurements == ref.measurements
    assert res.output == ref.output
    assert np.max(np.abs(out - ob.qs.a)) <= 1e-12


def test_program_with_fixed_noise_matches_oracle_backend():
    """Fixed(p) depolarizing noise after every gate: the same Philox draws on
    both sides."""
    import pyoracle as po
    from qoa_amd._native import NOISE_FIXED, QOA_CZ_REF_MASKS

    qs, res = P.run_text(PROGRAM, seed=9, noise_config=NoiseConfig.Fixed(0.2))
    with qs:
        out = qs.download()
    ob = OracleBackend(res.n_qubits, 9)
    ob.qs = po.QS(res.n_qubits, noise_kind=NOISE_FIXED, p=0.2, seed=9, flags=QOA_CZ_REF_MASKS)
    ref = P.execute(P.compile_text(PROGRAM).payload, ob)
    assert res.measurements == ref.measurements
    assert np.max(np.abs(out - ob.qs.a)) <= 1e-12
