"""The backend's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU only): tests/sanitize/driver.cpp drives the pass planner, the register
scheduler, the dense-block fuser and their host-only C ABI over random
circuits and checks the plans' invariants; any memory error or undefined
behaviour aborts the run."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent / "sanitize"


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                    reason="needs g++ and make")
def test_host_code_under_asan_ubsan():
    b = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    assert b.returncode == 0, b.stdout + b.stderr
    r = subprocess.run([str(HERE / "_build" / "host_asan"), "300"], capture_output=True,
                       text=True, timeout=600)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert "0 failures" in r.stdout
