#!/usr/bin/env python3
"""Gate-application benchmark on one MI355X.

Workload (default): the 28-qubit layered random circuit, depth 40, with CPHASE
ladders on every 5th layer (3164 gates), complex128 state resident in HBM
(4 GiB), FP64 arithmetic. One step = the whole circuit applied once.

Prints one JSON line: amp-updates/s (sum over gates of 2^n per second), the
roofline of the dominant kernel (algorithmic HBM bytes per launch / average
launch time, from HIP events on the state's stream) and a CPU baseline (the
CPU restatement of the reference path on a bounded sample of the same
circuit).

    python bench.py [--steps K] [--warmup W] [--method fused|per-gate]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(n: int, gates: np.ndarray, sample_gates: int, method_label: str) -> dict:
    """Times the CPU restatement (oracle, 1 thread) on the first gates of the
    same circuit and extrapolates per gate (the cost of a sweep does not depend
    on its position in the circuit)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po  # the checker's CPU path, timed as the baseline

    a = po.state(n)
    sub = np.ascontiguousarray(gates[:sample_gates])
    t0 = time.perf_counter()
    po.run_gates(a, sub)
    dt = time.perf_counter() - t0
    rate = len(sub) * float(1 << n) / dt
    return {"value": rate, "unit": "amp-updates/s", "cores": 1, "kind": "port",
            "sample": f"first {len(sub)} gates of the {n}-qubit circuit on the CPU "
                      f"restatement of src/vectorization.rs (1 thread, {dt:.1f} s), "
                      f"extrapolated per gate; host {os.cpu_count()} logical CPUs"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--method", default="auto", choices=["auto", "fused", "per-gate"])
    ap.add_argument("--qubits", type=int, default=28)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--seed", type=int, default=20241203)
    ap.add_argument("--cpu-sample-gates", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compare", action="store_true",
                    help="also time the other method and report it")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the state-vector path is "
              "single-device; multi-GPU sharding is a design note, not built)", file=sys.stderr)
        return 2

    from qoa_amd import StateVector
    from qoa_amd import workloads as W
    from qoa_amd._native import load

    lib = load()
    n = args.qubits
    gates = W.layered_random(n, args.depth, seed=args.seed, ladders=True)
    amp_updates = W.count_gate_amp_updates(n, gates)
    log(f"workload: {n} qubits, depth {args.depth}, {len(gates)} gates")

    method = args.method
    if method == "auto":
        method = "fused"
    sv = StateVector(n)

    def step(m: str) -> None:
        if m == "fused":
            sv.run_fused(gates)
        else:
            sv.run_gates(gates)

    try:
        step(method)
    except Exception as e:  # fused path unavailable in this build
        log(f"{method} unavailable ({e}); using per-gate")
        method = "per-gate"
    sv.apply_reset_all()
    for w in range(args.warmup):
        step(method)
        sv.synchronize()
        log(f"warmup {w + 1}/{args.warmup} done")

    sv.timing_enable(True)
    sv.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(method)
        log(f"step {k + 1}/{args.steps} enqueued")
    sv.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    kinds = {}
    for kind in ("sweep_dense", "sweep_diag", "sweep_perm", "fused_pass", "reduce", "other"):
        ms, cnt, b = sv.timing_read(kind)
        if cnt:
            kinds[kind] = {"ms": ms, "launches": cnt, "bytes": b,
                           "gbps": b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0,
                           "avg_us": ms * 1e3 / cnt}
    sv.timing_enable(False)
    dom = max(kinds.items(), key=lambda kv: kv[1]["ms"])
    tot_ms = sum(v["ms"] for v in kinds.values())
    tot_b = sum(v["bytes"] for v in kinds.values())
    ach = dom[1]["gbps"]
    value = amp_updates * args.steps / elapsed
    norm = sv.norm_sqr()
    log(f"{method}: {elapsed / args.steps * 1e3:.1f} ms/step, {value:.3e} amp-updates/s, "
        f"dominant {dom[0]} {ach:.0f} GB/s, norm^2 {norm:.15f}")

    other = None
    if args.compare:
        om = "per-gate" if method == "fused" else "fused"
        try:
            sv.synchronize()
            ta = time.perf_counter()
            step(om)
            sv.synchronize()
            other = {"method": om, "ms_per_step": (time.perf_counter() - ta) * 1e3,
                     "amp_updates_per_s": amp_updates / (time.perf_counter() - ta)}
        except Exception as e:
            other = {"method": om, "error": str(e)}

    name = ctypes_name(lib)
    sv.close()

    cpu = None
    if not args.no_cpu:
        log("CPU baseline sample ...")
        cpu = cpu_baseline(n, gates, args.cpu_sample_gates, method)
        log(f"CPU baseline {cpu['value']:.3e} amp-updates/s")

    out = {
        "metric": "amp-updates/s (28q FP64 random circuit)",
        "value": value,
        "unit": "amp-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{n}-qubit layered random circuit, depth {args.depth}, CPHASE "
                               f"ladders k=1..8 every 5th layer ({len(gates)} gates), complex128 "
                               "state resident in HBM", "n_qubits": n, "depth": args.depth,
                   "gates": len(gates), "seed": args.seed, "method": method},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": None, "kernel": dom[0],
                     "avg_launch_us": dom[1]["avg_us"],
                     "bytes_per_launch": dom[1]["bytes"] / dom[1]["launches"]},
        "hbm_gbps_all_kernels": tot_b / (tot_ms * 1e-3) / 1e9 if tot_ms else None,
        "kernels": kinds,
        "cpu_baseline": cpu,
        "device": name,
        "final_norm_sqr": norm,
    }
    if other:
        out["other_method"] = other
    print(json.dumps(out), flush=True)
    return 0


def ctypes_name(lib) -> str:
    import ctypes as C

    buf = C.create_string_buffer(256)
    hbm = C.c_uint64()
    cus = C.c_int()
    if lib.qoa_device_info(buf, 256, C.byref(hbm), C.byref(cus)) == 0:
        return f"{buf.value.decode()} {hbm.value / 2**30:.0f} GiB {cus.value} CUs"
    return "unknown"


if __name__ == "__main__":
    sys.exit(main())
