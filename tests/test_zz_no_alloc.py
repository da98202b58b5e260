"""The allocation-free sampler (qoa_sample_noalloc) through the C ABI against
qoa_sample on the same state and seed: two calls in a row, the second on a
smaller state with fewer samples, must give exactly the same outcomes; sizes
beyond its fixed scratch are refused."""
import ctypes as C

import numpy as np
import pytest

from kat_util import random_state
from qoa_amd import StateVector
from qoa_amd import _native as N

pytestmark = pytest.mark.gpu


def _call(lib, name, sv, seed, count):
    out = np.zeros(count, dtype=np.uint64)
    f = getattr(lib, name)
    f.argtypes, f.restype = N._SIGS[name]
    rc = f(sv.handle, seed, count, C.c_void_p(out.ctypes.data))
    return rc, out


def test_noalloc_sampler_equals_sampler_twice():
    lib = N.load()
    for n, count, seed in ((15, 3000, 11), (9, 700, 12)):
        with StateVector(n) as sv:
            sv.upload(random_state(n, seed=n))
            rc, want = _call(lib, "qoa_sample", sv, seed, count)
            assert rc == 0
            rc, got = _call(lib, "qoa_sample_noalloc", sv, seed, count)
            assert rc == 0, lib.qoa_last_error()
            assert np.array_equal(got, want), (n, count)
            assert np.all(got < (1 << n))


def test_noalloc_sampler_refuses_sizes_beyond_its_scratch():
    lib = N.load()
    with StateVector(6) as sv:
        rc, _ = _call(lib, "qoa_sample_noalloc", sv, 1, 65537)
        assert rc == N.QOA_ERR_ARG and b"scratch" in lib.qoa_last_error()
        rc, got = _call(lib, "qoa_sample_noalloc", sv, 1, 16)
        assert rc == 0 and np.all(got == 0)  # |000000> always samples 0
