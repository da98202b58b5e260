"""The C-ABI library loads and exports every function include/qoa_hip.h declares
(no compute call: this runs without a GPU)."""
import ctypes

from qoa_amd import _native
from qoa_amd import workloads as W


def test_header_symbols_exported(hip):
    names = _native.header_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(hip, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_ctypes_signatures_cover_header():
    names = set(_native.header_functions())
    assert names <= set(_native._SIGS), sorted(names - set(_native._SIGS))


def test_version_and_error_channel(hip):
    assert hip.qoa_version() >= 100
    # a null handle is rejected without touching the device
    rc = hip.qoa_apply_x(None, 0)
    assert rc == _native.QOA_ERR_HANDLE
    assert b"invalid handle" in hip.qoa_last_error()


def test_gate_struct_layout():
    assert ctypes.sizeof(_native.QoaGate) == 32
    assert _native.GATE_DTYPE.itemsize == 32


def test_workload_gate_counts():
    c1 = W.layered_random(16, 20)
    assert len(c1) == 16 * 20 + 10 * 8 + 10 * 7  # 320 one-qubit + 150 CNOT
    c2 = W.layered_random(28, 40, ladders=True)
    n1q, ncx = 28 * 40, 20 * 14 + 20 * 13
    assert len(c2) == n1q + ncx + 8 * 188
    ops = c2["op"]
    assert (ops == _native.OP["CPHASE"]).sum() == 1504
    assert (ops == _native.OP["CNOT"]).sum() == 540


def test_splitmix_matches_oracle():
    import pyoracle as po

    r = W.SplitMix64(1234567)
    assert [r.next() for _ in range(5)] == po.splitmix64(1234567, 5)
