import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP backend)")
    config.addinivalue_line("markers", "slow: long-running")


def _ensure_oracle():
    lib = ROOT / "oracle" / "_build" / "libqoa_oracle.so"
    if not lib.exists():
        sys.path.insert(0, str(ROOT / "oracle"))
        import build as obuild  # oracle/build.py

        obuild.build(verbose=False)


_ensure_oracle()


@pytest.fixture(scope="session")
def hip():
    """The HIP backend (libqoa_hip.so); GPU tests fail loudly without it."""
    from qoa_amd import _native

    return _native.load()
