"""Seeded synthetic circuits for the benchmark configurations.

All draws come from SplitMix64 (Steele, Lea & Flood 2014), so a (config,
seed) pair names one exact gate list. Angles are uniform in [0, 2*pi) as
((x >> 11) * 2^-53) * 2*pi.

  layered_random(n, depth, ladders)  16 q x 20 layers; 28 q x 40 layers with
      CPHASE ladders; 30 q x 60 layers. Each layer: one gate per qubit drawn
      uniformly from {H, RX, RY, RZ, PHASE}; then CNOT brickwork, pairs
      (0,1),(2,3).. on even layers and (1,2),(3,4).. on odd layers; with
      ladders, every layer L with L % 5 == 4 appends CPHASE(q, q+k, pi/2^k)
      for k = 1..8 and q = 0..n-1-k.
  ansatz_job(seed)  hardware-efficient ansatz: n ~ U{8..24}, depth ~ U{10..60},
      each layer RY(theta), RZ(phi) on every qubit then CZ(i, i+1).
  serving_job(seed) as ansatz_job but n in 8..26 with P(n) ~ 2^(-(n-8)/2).
"""
from __future__ import annotations

import math

import numpy as np

from ._native import GATE_DTYPE, OP

MASK64 = (1 << 64) - 1
TWO_PI = 2.0 * math.pi


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & MASK64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def angle(self) -> float:
        return self.uniform() * TWO_PI

    def below(self, k: int) -> int:
        return self.next() % k


_ONEQ = ("H", "RX", "RY", "RZ", "PHASE")


def layered_random(n: int, depth: int, seed: int = 20241203, ladders: bool = False) -> np.ndarray:
    rng = SplitMix64(seed)
    rows = []
    for layer in range(depth):
        for q in range(n):
            name = _ONEQ[rng.below(5)]
            th = 0.0 if name == "H" else rng.angle()
            rows.append((OP[name], q, 0, 0, th, 0))
        for q in range(layer % 2, n - 1, 2):
            rows.append((OP["CNOT"], q, q + 1, 0, 0.0, 0))
        if ladders and layer % 5 == 4:
            for k in range(1, 9):
                for q in range(0, n - k):
                    rows.append((OP["CPHASE"], q, q + k, 0, math.pi / (1 << k), 0))
    return np.array(rows, dtype=GATE_DTYPE)


def ansatz_gates(n: int, depth: int, rng: SplitMix64) -> np.ndarray:
    rows = []
    for _ in range(depth):
        for q in range(n):
            rows.append((OP["RY"], q, 0, 0, rng.angle(), 0))
            rows.append((OP["RZ"], q, 0, 0, rng.angle(), 0))
        for q in range(n - 1):
            rows.append((OP["CZ"], q, q + 1, 0, 0.0, 0))
    return np.array(rows, dtype=GATE_DTYPE)


def ansatz_job(seed: int, n_min: int = 8, n_max: int = 24, d_min: int = 10, d_max: int = 60):
    rng = SplitMix64(seed)
    n = n_min + rng.below(n_max - n_min + 1)
    depth = d_min + rng.below(d_max - d_min + 1)
    return n, depth, ansatz_gates(n, depth, rng)


def serving_qubits(rng: SplitMix64, n_min: int = 8, n_max: int = 26) -> int:
    w = [2.0 ** (-(k - n_min) / 2.0) for k in range(n_min, n_max + 1)]
    tot = sum(w)
    u = rng.uniform() * tot
    acc = 0.0
    for k, wk in zip(range(n_min, n_max + 1), w):
        acc += wk
        if u < acc:
            return k
    return n_max


def serving_job(seed: int):
    rng = SplitMix64(seed)
    n = serving_qubits(rng)
    depth = 10 + rng.below(51)
    return n, depth, ansatz_gates(n, depth, rng)


def count_gate_amp_updates(n: int, gates: np.ndarray) -> float:
    """amp-updates = sum over gates of 2^n."""
    return float(len(gates)) * float(1 << n)


def algorithmic_bytes(n: int, gates: np.ndarray) -> float:
    """Per-gate sweep bytes of the per-gate method (read + write of the lines
    a gate touches, 64-byte lines of 4 complex128 amplitudes)."""
    N = float(1 << n)
    tot = 0.0
    for g in gates:
        op = int(g["op"])
        q0, q1 = int(g["q0"]), int(g["q1"])
        if op in (OP["Z"], OP["S"], OP["T"], OP["PHASE"], OP["PHASESHIFT"]):
            tot += 16 * N if q0 >= 2 else 32 * N
        elif op == OP["CNOT"]:
            tot += 16 * N if q0 >= 2 else 32 * N
        elif op in (OP["CZ"], OP["CPHASE"]):
            k = (q0 >= 2) + (q1 >= 2)
            tot += 32 * N / (1 << k)
        else:
            tot += 32 * N
    return tot
