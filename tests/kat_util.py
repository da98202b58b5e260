"""Shared helpers: reference KAT fixtures and random states."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def kats():
    return json.loads((GOLDEN / "kats.json").read_text())["cases"]


def arr(pairs) -> np.ndarray:
    a = np.array([complex(r, i) for r, i in pairs], dtype=np.complex128)
    return np.ascontiguousarray(a)


def random_state(n: int, seed: int, normalise: bool = True) -> np.ndarray:
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    if normalise:
        a /= np.linalg.norm(a)
    return np.ascontiguousarray(a.astype(np.complex128))


def assert_kat(actual: np.ndarray, expected: np.ndarray, eps: float, name: str) -> None:
    """src/test.rs:33-59: re and im each within eps."""
    assert actual.shape == expected.shape, name
    dr = np.abs(actual.real - expected.real)
    di = np.abs(actual.imag - expected.imag)
    assert (dr < eps).all() and (di < eps).all(), f"{name}: max diff {max(dr.max(), di.max())}"
