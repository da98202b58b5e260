"""Specialised passes on the CPU: the generated per-pass kernels of several
plans compile for gfx950 (hipRTC, no device needed), for the exact and the
FMA/merged arithmetic and both index widths. The GPU tests compare their
results with the op-loop kernel's."""
import ctypes as C

import numpy as np
import pytest

from qoa_amd import _native as N
from qoa_amd import workloads as W
from qoa_amd.state import gates_array
from test_planner import OP, mixed_circuit


def _compile(n, g, flags, mats=None, tile=0):
    lib = N.load()
    ga = gates_array(g)
    m = None if mats is None else np.ascontiguousarray(mats, dtype=np.complex128)
    opts = N.QoaFuseOpts(tile, 0, flags | N.QOA_FUSE_SPECIALIZE, 0)
    return lib.qoa_specialize_compile(n, N.ptr(ga), len(ga), None if m is None else N.ptr(m),
                                      C.byref(opts), b"gfx950")


def _info():
    lib = N.load()
    p, c, t = C.c_uint32(), C.c_uint32(), C.c_double()
    assert lib.qoa_specialize_info(C.byref(p), C.byref(c), C.byref(t)) == 0
    return p.value, c.value, t.value


@pytest.fixture(autouse=True)
def _private_code_cache(tmp_path, monkeypatch):
    monkeypatch.setenv("QOA_SPEC_CACHE", str(tmp_path / "spec"))


@pytest.mark.parametrize("flags", [0, N.QOA_FUSE_MERGE_DIAG, N.QOA_FUSE_FACTOR | N.QOA_FUSE_MERGE_DIAG])
def test_layered_circuit_compiles(flags, tmp_path):
    g = W.layered_random(14, 6, seed=3 + flags, ladders=True)
    before = _info()
    assert _compile(14, g, flags) == 0, N.load().qoa_last_error()
    after = _info()
    assert after[1] == before[1] + 1 and after[2] > before[2]
    assert len(list((tmp_path / "spec").glob("*.co"))) == 1  # the on-disk copy
    # the same structure again comes from the process-wide cache
    assert _compile(14, g, flags) == 0
    assert _info()[1] == after[1]


def test_mixed_surface_compiles():
    """Every gate kind the register passes take (2x2 and 4x4 matrices, Y,
    SWAP, CSWAP, Toffoli, controlled ops on register and thread bits), exact
    arithmetic, 8-qubit tiles."""
    g, mats = mixed_circuit(11, 120, seed=21)
    assert _compile(11, g, 0, mats, tile=8) == 0, N.load().qoa_last_error()
    assert _compile(11, g, N.QOA_FUSE_FMA, mats, tile=8) == 0, N.load().qoa_last_error()


def test_rejects_missing_arch():
    lib = N.load()
    g = gates_array(W.layered_random(10, 2, seed=1))
    assert lib.qoa_specialize_compile(10, N.ptr(g), len(g), None, None, b"") == N.QOA_ERR_ARG
