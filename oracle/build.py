"""Builds the CPU oracle (test infrastructure only) into oracle/_build/.

    python oracle/build.py

The oracle is compiled with gcc -O3 -ffp-contract=off and no fast-math,
mirroring the reference's release profile (Cargo.toml:94-103): no FMA
contraction, so each gate's arithmetic matches the GPU kernels bit for bit.

The reference itself (Rust, nightly, 416 crates from crates.io) cannot be
built here, so oracle/_ref/ stays empty.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build"
LIB = OUT / "libqoa_oracle.so"

FLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
         "-std=c11", "-Wall", "-D_DEFAULT_SOURCE"]


def build(verbose: bool = True) -> Path:
    OUT.mkdir(exist_ok=True)
    tmp = OUT / "libqoa_oracle.so.tmp"
    cmd = ["gcc", *FLAGS, str(HERE / "qoa_oracle.c"), "-o", str(tmp), "-lm", "-fopenmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", flush=True)
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
