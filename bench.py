#!/usr/bin/env python3
"""Gate-application benchmark on one MI355X.

Headline (metric/value): the 28-qubit layered random circuit, depth 40, with
CPHASE ladders on every 5th layer (3164 gates), complex128 state resident in
HBM (4 GiB), FP64 arithmetic, run with fused register-resident passes. One
step = the whole circuit applied once. value = amp-updates/s (sum over gates
of 2^n per second).

Also reported in the same JSON line:
  roofline      dominant kernel: algorithmic HBM bytes per launch / average
                launch time (HIP events on the state's stream), vs 8 TB/s
  cpu_baseline  the CPU restatement of the reference loops on a bounded
                sample of the same circuit (1 thread)
  extras        per-gate method on the same circuit; 16-qubit circuit
                latency; 1024-job batched ansatz (jobs/s); a Poisson serving
                run (p50/p99 latency, jobs/s)

    python bench.py [--steps K] [--warmup W] [--no-extras] [--serve-rates 500,2000]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_MFMA_TFLOPS = 157.3
FP64_TFLOPS = 78.6
SIMDS = 1024          # 256 CUs x 4 SIMDs; a wave64 VALU or SALU instruction issues in 4 cycles
CLOCK_MHZ = 2400.0    # peak engine clock
QOA_FUSE_LDS_ONLY, QOA_FUSE_FMA = 2, 4


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(n: int, gates: np.ndarray, sample_gates: int) -> dict:
    """Times the CPU restatement of the reference loops on the first gates of
    the same circuit (the cost of a sweep does not depend on its position):
      strong   - every gate loop split over all host cores, no renormalisation
                 (the headline value);
      faithful - the reference's QuantumState model: one-thread sweep plus an
                 all-core normalize after every gate
                 (src/vectorization.rs:926-938, src/runtime/quantum_state.rs:868-900);
      sweep_1t - the one-thread sweeps alone."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po  # the checker's CPU path, timed as the baseline

    cores = max(1, min(os.cpu_count() or 1, 16))

    def timed(fn, count):
        a = po.state(n)
        sub = np.ascontiguousarray(gates[:count])
        t0 = time.perf_counter()
        fn(a, sub)
        dt = time.perf_counter() - t0
        return count * float(1 << n) / dt, dt

    po.set_threads(cores)
    try:
        strong, t_strong = timed(po.run_gates, sample_gates * 8)
    finally:
        po.set_threads(1)
    faithful, t_faith = timed(lambda a, s: po.run_gates_faithful(a, s, cores), sample_gates)
    one, t_one = timed(po.run_gates, sample_gates)
    return {"value": strong, "unit": "amp-updates/s", "cores": cores, "kind": "port",
            "sample": f"first {sample_gates * 8} gates of the {n}-qubit circuit on the CPU "
                      f"restatement of src/vectorization.rs with every gate loop split over "
                      f"{cores} threads ({t_strong:.1f} s), extrapolated per gate; host has "
                      f"{os.cpu_count()} logical CPUs",
            "variants": {
                "faithful": {"value": faithful, "threads": cores,
                             "sample": f"first {sample_gates} gates, one-thread sweep + "
                                       f"{cores}-thread normalize per gate ({t_faith:.1f} s)"},
                "sweep_1thread": {"value": one, "threads": 1,
                                  "sample": f"first {sample_gates} gates, one-thread sweeps "
                                            f"({t_one:.1f} s)"}}}


def kernel_table(sv) -> dict:
    kinds = {}
    for kind in ("sweep_dense", "sweep_diag", "sweep_perm", "fused_pass", "fused_block",
                 "reduce", "other"):
        ms, cnt, b = sv.timing_read(kind)
        if cnt:
            kinds[kind] = {"ms": ms, "launches": cnt, "bytes": b,
                           "gbps": b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0,
                           "avg_us": ms * 1e3 / cnt}
    return kinds


def run_circuit(sv, gates, method: str, steps: int, warmup: int, tag: str):
    def step():
        if method == "per-gate":
            sv.run_gates(gates)
        elif method == "fused-lds":
            sv.run_fused(gates, flags=QOA_FUSE_LDS_ONLY)
        elif method == "fused-fma":
            sv.run_fused(gates, flags=QOA_FUSE_FMA)
        else:
            sv.run_fused(gates)

    sv.apply_reset_all()
    for w in range(warmup):
        step()
        sv.synchronize()
        log(f"{tag}: warmup {w + 1}/{warmup} done")
    sv.timing_enable(True)
    sv.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step()
        log(f"{tag}: step {k + 1}/{steps} enqueued")
    sv.synchronize()
    elapsed = time.perf_counter() - t0
    kinds = kernel_table(sv)
    sv.timing_enable(False)
    return elapsed, kinds


def extra_c1(args) -> dict:
    from qoa_amd import StateVector
    from qoa_amd import workloads as W

    n = 16
    g = W.layered_random(n, 20, seed=args.seed)
    out = {"workload": "16-qubit layered random circuit, depth 20 (470 gates)"}
    with StateVector(n) as sv:
        for method in ("per-gate", "fused"):
            el, _ = run_circuit(sv, g, method, 20, 2, f"c1 {method}")
            out[f"{method}_us_per_circuit"] = el / 20 * 1e6
            out[f"{method}_circuits_per_s"] = 20 / el
    return out


def extra_c3(args) -> dict:
    """30 qubits, depth 60 with CPHASE ladders: dense 5-qubit blocks on FP32
    matrix cores (complex64, 8 GiB) against the unfused FP64 per-gate run
    (complex128, 16 GiB), both resident, compared on the device."""
    from qoa_amd import StateVector
    from qoa_amd import workloads as W
    from qoa_amd._native import QOA_C64

    n, depth = 30, 60
    g = W.layered_random(n, depth, seed=args.seed, ladders=True)
    out = {"workload": f"{n}-qubit layered random circuit, depth {depth}, CPHASE ladders "
                       f"({len(g)} gates)"}
    with StateVector(n, precision=QOA_C64) as s32:
        s32.timing_enable(True)
        t0 = time.perf_counter()
        nb = s32.run_blocks(g)
        s32.synchronize()
        out["fp32_blocks_s"] = time.perf_counter() - t0
        kk = kernel_table(s32)
        d = f = float("nan")
        if not args.c3_skip_ref:
            with StateVector(n) as ref:
                t0 = time.perf_counter()
                ref.run_gates(g)
                ref.synchronize()
                out["fp64_per_gate_s"] = time.perf_counter() - t0
                log("c3: FP64 reference done")
                d, f = ref.compare(s32)
    blk = kk["fused_block"]
    flops = float(2 ** (n + 8))
    out.update({"blocks": nb, "gates_per_block": len(g) / nb, "max_abs_diff": d,
                "one_minus_fidelity": 1.0 - f, "block_avg_us": blk["avg_us"],
                "block_hbm_gbps": blk["gbps"], "block_hbm_frac": blk["gbps"] / HBM_PEAK_GBS,
                "block_tflops": flops / (blk["avg_us"] * 1e-6) / 1e12,
                "block_mfma_frac": flops / (blk["avg_us"] * 1e-6) / 1e12 / FP32_MFMA_TFLOPS,
                "kernels": kk})
    return out


def extra_c4(args) -> dict:
    from qoa_amd import workloads as W
    from qoa_amd.batch import Batch

    t0 = time.perf_counter()
    jobs = [W.ansatz_job(args.seed + j) for j in range(1024)]
    gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    b = Batch([(n, g) for n, _, g in jobs], tile_qubits=12)
    plan_s = time.perf_counter() - t0
    info = b.info()
    b.run()
    b.synchronize()  # warm-up
    times = []
    for _ in range(3):
        b.run()
        times.append(b.synchronize())
    ms = min(times)
    b.close()
    amp_updates = float(sum(len(g) * (1 << n) for n, _, g in jobs))
    return {"workload": "1024 hardware-efficient-ansatz circuits, n~U{8..24}, depth~U{10..60}, "
                        "FP64, all states resident", "jobs": 1024,
            "state_gib": info["total_amps"] * 16 / 2**30, "rounds": info["rounds"],
            "ms_per_batch": ms, "jobs_per_s": 1024 / (ms * 1e-3),
            "amp_updates_per_s": amp_updates / (ms * 1e-3),
            "hbm_gbps": info["alg_bytes"] / (ms * 1e-3) / 1e9,
            "gen_s": gen, "plan_s": plan_s}


def extra_c5(args, rate: float, n_jobs: int) -> dict:
    from qoa_amd import workloads as W
    from qoa_amd.server import Server

    rng = np.random.default_rng(args.seed)
    jobs = [W.serving_job(args.seed * 7 + j) for j in range(n_jobs)]
    gaps = rng.exponential(1.0 / rate, size=n_jobs)
    with Server(budget_bytes=int(200 * 2**30)) as srv:
        t_start = srv.now() + 0.2
        arrivals = t_start + np.cumsum(gaps)
        ids = [srv.submit(n, g, trajectories=64, p1=1e-3, p2=1e-2, noise_seed=args.seed + j,
                          arrival=float(arrivals[j]))
               for j, (n, _, g) in enumerate(jobs)]
        res = [srv.wait(i, timeout=600) for i in ids]
        stats = srv.stats()
    ok = [r for r in res if r.status == 0]
    lat = np.array([r.latency for r in ok])
    span = max(r.t_done for r in ok) - t_start
    return {"rate_jobs_per_s": rate, "jobs": n_jobs, "completed": len(ok),
            "rejected": stats["rejected"], "jobs_per_s": len(ok) / span,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "steps": stats["steps"], "arena_peak_gib": stats["arena_peak"] / 2**30}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--method", default="fused",
                    choices=["fused", "fused-fma", "fused-lds", "per-gate"])
    ap.add_argument("--qubits", type=int, default=28)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--seed", type=int, default=20241203)
    ap.add_argument("--cpu-sample-gates", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--extras", default="",
                    help="comma list of extras to run (c1,c2,c3,c4,c5; default all)")
    ap.add_argument("--serve-rates", default="1000")
    ap.add_argument("--serve-jobs", type=int, default=1500)
    ap.add_argument("--c3-skip-ref", action="store_true",
                    help="C3: time the FP32 blocks only (no FP64 reference run, no comparison)")
    ap.add_argument("--calibrate", action="store_true",
                    help="only time the copy-like X sweep (reads and writes every amplitude "
                         "once) used to scale HBM counter readings")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the state-vector path is "
              "single-device; multi-GPU sharding is a design note, not built)", file=sys.stderr)
        return 2
    if args.calibrate:
        return calibrate(args)

    from qoa_amd import StateVector
    from qoa_amd import workloads as W
    from qoa_amd._native import load

    lib = load()
    n = args.qubits
    gates = W.layered_random(n, args.depth, seed=args.seed, ladders=True)
    amp_updates = W.count_gate_amp_updates(n, gates)
    log(f"workload: {n} qubits, depth {args.depth}, {len(gates)} gates")

    sv = StateVector(n)
    elapsed, kinds = run_circuit(sv, gates, args.method, args.steps, args.warmup, "c2")
    dom = max(kinds.items(), key=lambda kv: kv[1]["ms"])
    tot_ms = sum(v["ms"] for v in kinds.values())
    tot_b = sum(v["bytes"] for v in kinds.values())
    value = amp_updates * args.steps / elapsed
    norm = sv.norm_sqr()
    log(f"{args.method}: {elapsed / args.steps * 1e3:.1f} ms/step, {value:.3e} amp-updates/s, "
        f"dominant {dom[0]} {dom[1]['gbps']:.0f} GB/s, norm^2 {norm:.15f}")

    want = set(x for x in args.extras.split(",") if x) if args.extras else None
    if args.no_extras:
        want = set()

    def wanted(label: str) -> bool:
        return want is None or label in want

    extras = {}
    if wanted("c2"):
        for other in ("fused", "fused-fma", "per-gate", "fused-lds"):
            if other == args.method:
                continue
            el, kk = run_circuit(sv, gates, other, 1, 0, f"c2 {other}")
            extras[f"c2_{other}"] = {"ms_per_step": el * 1e3,
                                     "amp_updates_per_s": amp_updates / el,
                                     "kernels": kk}
    name = device_name(lib)
    sv.close()

    for label, fn in (("c1", lambda: extra_c1(args)), ("c3_blocks", lambda: extra_c3(args)),
                      ("c4_batch", lambda: extra_c4(args))):
        if not wanted(label.split("_")[0]):
            continue
        try:
            extras[label] = fn()
            log(f"{label}: {json.dumps(extras[label])[:200]}")
        except Exception as e:  # noqa: BLE001
            extras[label] = {"error": str(e)}
            log(f"{label} failed: {e}")
    if wanted("c5"):
        serve = []
        for r in [float(x) for x in args.serve_rates.split(",") if x]:
            try:
                serve.append(extra_c5(args, r, args.serve_jobs))
                log(f"c5 serving: {serve[-1]}")
            except Exception as e:  # noqa: BLE001
                serve.append({"rate_jobs_per_s": r, "error": str(e)})
                log(f"c5 serving failed: {e}")
        extras["c5_serving"] = serve

    cpu = None
    if not args.no_cpu:
        log("CPU baseline sample ...")
        cpu = cpu_baseline(n, gates, args.cpu_sample_gates)
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
                   "gates": len(gates), "seed": args.seed, "method": args.method},
        "roofline": roofline(dom, args),
        "hbm_gbps_all_kernels": tot_b / (tot_ms * 1e-3) / 1e9 if tot_ms else None,
        "kernels": kinds,
        "cpu_baseline": cpu,
        "extras": extras,
        "device": name,
        "final_norm_sqr": norm,
    }
    print(json.dumps(out), flush=True)
    return 0


def measured_profile(key: str):
    """PMC figures committed under profiles/ for this kernel and workload."""
    p = ROOT / "profiles" / "traffic.json"
    if not p.exists():
        return None
    return json.loads(p.read_text()).get(key)


def roofline(dom, args) -> dict:
    kind, k = dom
    per_launch = k["bytes"] / k["launches"]
    out = {"bound": "hbm", "achieved": k["gbps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": k["gbps"] / HBM_PEAK_GBS, "traffic": None, "kernel": kind,
           "avg_launch_us": k["avg_us"], "bytes_per_launch": per_launch}
    default = (args.qubits, args.depth, args.seed) == (28, 40, 20241203)
    prof = measured_profile("fused_pass/28q-d40-ladders") if (
        default and kind == "fused_pass" and args.method == "fused") else None
    if prof:
        out["traffic"] = prof["bytes_per_launch"]
        out["traffic_source"] = "profiles/traffic.json (rocprofv3 FETCH_SIZE/WRITE_SIZE, " \
                                "copy-calibrated)"
        # the pass is bound by instruction issue, not HBM: its FP64 rate and the
        # share of the pass's cycles each SIMD spends issuing VALU / SALU work
        fl = prof["fp64_valu_instr"] * 64 / (k["avg_us"] * 1e-6) / 1e12
        out["fp64_valu"] = {"achieved_tflops": fl, "peak_tflops": FP64_TFLOPS,
                            "frac": fl / FP64_TFLOPS,
                            "fp64_share_of_valu_instr": prof["fp64_valu_instr"] / prof["valu_instr"]}
        cycles = k["avg_us"] * CLOCK_MHZ
        out["issue"] = {"valu_busy": prof["valu_instr"] / SIMDS * 4 / cycles,
                        "salu_busy": prof["salu_instr"] / SIMDS * 4 / cycles}
    return out


def calibrate(args) -> int:
    """X on the top qubit: a pure permutation sweep, 2^n * 16 B read and
    written per launch (the copy kernel of known size)."""
    from qoa_amd import StateVector

    n = args.qubits
    with StateVector(n) as sv:
        sv.apply_reset_all()
        sv.apply_x(n - 1)
        sv.synchronize()
        sv.timing_enable(True)
        for _ in range(args.steps):
            sv.apply_x(n - 1)
        sv.synchronize()
        kinds = kernel_table(sv)
    perm = kinds["sweep_perm"]
    print(json.dumps({"calibration": f"X sweep on qubit {n - 1} of a {n}-qubit complex128 state",
                      "launches": perm["launches"], "avg_us": perm["avg_us"],
                      "read_bytes_per_launch": 16 * 2**n, "write_bytes_per_launch": 16 * 2**n,
                      "gbps": perm["gbps"]}), flush=True)
    return 0


def device_name(lib) -> str:
    import ctypes as C

    buf = C.create_string_buffer(256)
    hbm = C.c_uint64()
    cus = C.c_int()
    if lib.qoa_device_info(buf, 256, C.byref(hbm), C.byref(cus)) == 0:
        return f"{buf.value.decode().strip()} {hbm.value / 2**30:.0f} GiB {cus.value} CUs"
    return "unknown"


if __name__ == "__main__":
    sys.exit(main())
