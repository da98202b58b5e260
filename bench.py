#!/usr/bin/env python3
"""Gate-application benchmark on one MI355X.

Headline (metric/value): the 28-qubit layered random circuit, depth 40, with
CPHASE ladders on every 5th layer (3164 gates), complex128 state resident in
HBM (4 GiB), FP64 arithmetic, run with fused register-resident passes. One
step = the whole circuit applied once. value = amp-updates/s (sum over gates
of 2^n per second).

Also reported in the same JSON line:
  accuracy      max |d amp| and 1 - fidelity of the headline method against
                the per-gate FP64 sweeps of the same circuit (which equal the
                CPU oracle bit for bit at this size)
  roofline      dominant kernel: algorithmic HBM bytes per launch / average
                launch time (HIP events on the state's stream), vs 8 TB/s, and
                the copy ceiling measured on the same state
  cpu_baseline  the CPU restatement of the reference loops on a bounded
                sample of the same circuit
  extras        the other methods on the same circuit (with their errors);
                16-qubit circuit latency; 30-qubit FP32 blocks; 1024-job
                batched ansatz (FP64 jobs/s vs the plain per-gate per-job FP64
                loop); Poisson serving swept over arrival rates (p50/p99
                latency, jobs/s, vs the plain per-gate trajectory loop); the
                reference's per-gate timed units (benches/benchmark.rs)

    python bench.py [--steps K] [--warmup W] [--no-extras] [--extras c1,c2,...]
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
QOA_FUSE_LDS_ONLY, QOA_FUSE_FMA, QOA_FUSE_MERGE_DIAG, QOA_FUSE_SPECIALIZE = 2, 4, 16, 64
QOA_FUSE_FACTOR, QOA_FUSE_RELABEL = 128, 256


METHOD_NOTES = {
    "fused-relabel": "fused register passes (12-qubit tiles) as kernels compiled for this "
                     "circuit (hipRTC, before the timed steps; extras.c2_specialize), FP64 with "
                     "fused multiply-adds, runs of diagonal gates merged per register group, "
                     "scalar factors of H/RX/RY/RZ taken out and applied once per pass, qubits "
                     "relabelled between passes (each pass stores the next pass's qubits at the "
                     "contiguous low positions; the order is restored before the run returns): "
                     "within 1e-12 of the per-gate FP64 sweeps (accuracy field)",
    "fused-factor": "fused register passes (12-qubit tiles) as kernels compiled for this "
                    "circuit (hipRTC, before the timed steps; extras.c2_specialize), FP64 with "
                    "fused multiply-adds, runs of diagonal gates merged per register group, "
                    "scalar factors of H/RX/RY/RZ taken out and applied once per pass; within "
                    "1e-12 of the per-gate FP64 sweeps (accuracy field)",
    "fused-spec": "fused register passes (12-qubit tiles) as kernels compiled for this circuit "
                  "(hipRTC, before the timed steps; extras.c2_specialize), FP64 with fused "
                  "multiply-adds, runs of diagonal gates merged per register group: the "
                  "fused-merge op-loop kernel's results bit for bit, within 1e-12 of the "
                  "per-gate FP64 sweeps (accuracy field)",
    "fused-merge": "fused register passes (12-qubit tiles), FP64 with fused multiply-adds, "
                   "runs of diagonal gates merged per register group; within 1e-12 of the "
                   "per-gate FP64 sweeps (accuracy field)",
    "fused": "fused register passes, FP64 without FMA: bit-identical to the oracle in the "
             "planned order",
    "fused-fma": "fused register passes, FP64 with fused multiply-adds",
    "fused-lds": "fused passes with the tile in LDS",
    "per-gate": "one sweep kernel per gate (bit-identical to the oracle)",
}


def specialize_info(lib) -> dict:
    import ctypes as C
    p, c, t = C.c_uint32(), C.c_uint32(), C.c_double()
    lib.qoa_specialize_info(C.byref(p), C.byref(c), C.byref(t))
    return {"programs": p.value, "compiles": c.value, "compile_seconds": t.value}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    """Threads for the all-core CPU legs: the OpenMP thread count the host
    is set up for (OMP_NUM_THREADS), else every CPU this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def strided(gates: np.ndarray, count: int) -> np.ndarray:
    """`count` gates spread evenly over the circuit (every kind and qubit
    range of it, not just its first layer)."""
    idx = np.linspace(0, len(gates) - 1, num=min(count, len(gates))).astype(np.int64)
    return np.ascontiguousarray(gates[idx])


def cpu_baseline(n: int, gates: np.ndarray, sample_gates: int) -> dict:
    """Times the CPU restatement of the reference loops on gates spread over
    the same circuit (a sweep's cost depends on its kind and qubits, not on
    its position, so the rate extrapolates per gate):
      strong   - every gate loop split over all host cores, no renormalisation
                 (the headline value);
      faithful - the reference's QuantumState model: one-thread sweep plus an
                 all-core normalize after every gate
                 (src/vectorization.rs:926-938, src/runtime/quantum_state.rs:868-900);
      sweep_1t - the one-thread sweeps alone.
    Also the reference's timed units: each gate on a fresh |0...0> (allocation
    included, benches/benchmark.rs:91-252) and the cpu-bench loop of H with a
    normalize per gate (src/main.rs:1260-1311)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po  # the checker's CPU path, timed as the baseline

    cores = host_cores()

    def timed(fn, sub):
        a = po.state(n)
        t0 = time.perf_counter()
        fn(a, sub)
        dt = time.perf_counter() - t0
        return len(sub) * float(1 << n) / dt, dt

    strong_gates = strided(gates, sample_gates * 8)
    few = strided(gates, sample_gates)
    po.set_threads(cores)
    try:
        strong, t_strong = timed(po.run_gates, strong_gates)
    finally:
        po.set_threads(1)
    faithful, t_faith = timed(lambda a, sub: po.run_gates_faithful(a, sub, cores), few)
    one, t_one = timed(po.run_gates, few)
    units = cpu_units(po)
    return {"value": strong, "unit": "amp-updates/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "logical_cpus": os.cpu_count(),
            "sample": f"{len(strong_gates)} gates spread evenly over the {n}-qubit circuit on "
                      f"the CPU restatement of src/vectorization.rs, every gate loop split over "
                      f"{cores} threads ({t_strong:.1f} s), extrapolated per gate",
            "variants": {
                "faithful": {"value": faithful, "threads": cores,
                             "sample": f"{len(few)} gates spread over the circuit, one-thread "
                                       f"sweep + {cores}-thread normalize per gate "
                                       f"({t_faith:.1f} s)"},
                "sweep_1thread": {"value": one, "threads": 1,
                                  "sample": f"{len(few)} gates spread over the circuit, "
                                            f"one-thread sweeps ({t_one:.1f} s)"}},
            "units": units}


A20_GATES = ("hadamard", "rx", "ry", "rz", "phaseshift", "reset", "reset_all", "t", "s",
             "cnot", "swap", "cz", "cphase", "cswap")


def cpu_units(po) -> dict:
    """benches/benchmark.rs:91-252 on the CPU restatement (one thread, as the
    reference's loops): a fresh |0...0> state and one gate per iteration,
    target 0 (controls/targets 0 and 1, CSWAP(2, 0, 1), CZ/CPHASE masks 0
    and 1 as the bench passes them); and cpu-bench (src/main.rs:1260-1311):
    H on qubit 0 with a normalize after each (QuantumState::apply_h)."""
    out = {}
    for n in (16, 20):
        reps = 40 if n == 16 else 6
        row = {}
        for name in A20_GATES:
            t0 = time.perf_counter()
            for _ in range(reps):
                a = po.state(n)
                cpu_unit_gate(po, a, name)
            row[name] = (time.perf_counter() - t0) / reps * 1e6
        out[f"{n}q_us_per_iter"] = row
    hb = {}
    for n in (20, 24):
        qs = po.QS(n)
        g = np.array([(0x05, 0, 0, 0, 0.0, 0)], dtype=po_gate_dtype())
        ops = 5
        t0 = time.perf_counter()
        for _ in range(ops):
            qs.apply(g)
        hb[f"{n}q_ops_per_s"] = ops / (time.perf_counter() - t0)
    out["cpu_bench_h_normalize"] = hb
    return out


def po_gate_dtype():
    from qoa_amd._native import GATE_DTYPE

    return GATE_DTYPE


def cpu_unit_gate(po, a, name: str) -> None:
    o = po.orc
    if name == "hadamard":
        o.hadamard(a, 1 / math.sqrt(2), 0.0, 0)
    elif name in ("rx", "ry", "rz"):
        getattr(o, name)(a, 0, math.pi / 2)
    elif name == "phaseshift":
        o.phaseshift(a, 0, math.pi / 3)
    elif name == "reset":
        o.reset(a, 0)
    elif name == "reset_all":
        o.reset_all(a)
    elif name in ("t", "s"):
        getattr(o, name)(a, 0)
    elif name == "cnot":
        o.cnot(a, 0, 1)
    elif name == "swap":
        o.swap(a, 0, 1)
    elif name == "cz":
        o.cz(a, 0, 1)
    elif name == "cphase":
        o.cphase(a, 0, 1, math.pi / 4)
    elif name == "cswap":
        o.controlled_swap(a, 2, 0, 1)


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


def run_circuit(sv, gates, method: str, steps: int, warmup: int, tag: str,
                timing: bool = True):
    """Times `steps` back-to-back runs (after `warmup` waited-for ones); with
    timing, HIP events around every kernel feed the per-kind table."""
    def step():
        if method == "per-gate":
            sv.run_gates(gates)
        elif method == "fused-lds":
            sv.run_fused(gates, flags=QOA_FUSE_LDS_ONLY)
        elif method == "fused-fma":
            sv.run_fused(gates, flags=QOA_FUSE_FMA)
        elif method == "fused-merge":
            sv.run_fused(gates, flags=QOA_FUSE_MERGE_DIAG)
        elif method == "fused-spec":
            sv.run_fused(gates, flags=QOA_FUSE_MERGE_DIAG | QOA_FUSE_SPECIALIZE)
        elif method == "fused-factor":
            sv.run_fused(gates, flags=QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG | QOA_FUSE_SPECIALIZE)
        elif method == "fused-relabel":
            sv.run_fused(gates, flags=QOA_FUSE_RELABEL | QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG
                         | QOA_FUSE_SPECIALIZE)
        else:
            sv.run_fused(gates)

    sv.apply_reset_all()
    for w in range(warmup):
        step()
        sv.synchronize()
        log(f"{tag}: warmup {w + 1}/{warmup} done")
    sv.timing_enable(timing)
    sv.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step()
        if timing:
            log(f"{tag}: step {k + 1}/{steps} enqueued")
    sv.synchronize()
    elapsed = time.perf_counter() - t0
    kinds = kernel_table(sv) if timing else {}
    sv.timing_enable(False)
    return elapsed, kinds


def extra_c1(args) -> dict:
    from qoa_amd import StateVector
    from qoa_amd import workloads as W

    n = 16
    g = W.layered_random(n, 20, seed=args.seed)
    out = {"workload": "16-qubit layered random circuit, depth 20 (470 gates)",
           "note": "fused: exact op-loop kernel (bit-identical to the oracle in planned order); "
                   "fused-relabel: the headline's mode (specialised, factored, relabelled; "
                   "compiled in the warm-up calls), compared with the per-gate run"}
    with StateVector(n) as sv, StateVector(n) as ref:
        # no per-kernel events here: at 16 qubits they would cost as much as the kernels
        el, _ = run_circuit(ref, g, "per-gate", 20, 2, "c1 per-gate", timing=False)
        out["per-gate_us_per_circuit"] = el / 20 * 1e6
        out["per-gate_circuits_per_s"] = 20 / el
        ref.apply_reset_all()
        ref.run_gates(g)
        for method in ("fused", "fused-relabel"):
            el, _ = run_circuit(sv, g, method, 20, 2, f"c1 {method}", timing=False)
            out[f"{method}_us_per_circuit"] = el / 20 * 1e6
            out[f"{method}_circuits_per_s"] = 20 / el
            sv.apply_reset_all()
            run_circuit(sv, g, method, 1, 0, f"c1 {method} check")
            d, f = sv.compare(ref)
            out[f"{method}_max_abs_diff_vs_per_gate"] = d
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


def compare_host(a: np.ndarray, ref: np.ndarray) -> tuple[float, float]:
    d = float(np.max(np.abs(a - ref)))
    f = abs(np.vdot(ref, a)) ** 2 / (np.vdot(ref, ref).real * np.vdot(a, a).real)
    return d, float(1.0 - f)


def plain_c4_loop(jobs) -> tuple[float, dict]:
    """The plain per-gate, per-job FP64 GPU loop: each job from |0...0> with
    one sweep kernel per gate (qoa_run_gates), one job after the other (the
    loop waits for each job's result). One state handle per qubit count,
    allocated before timing. Returns the loop's wall time and the final
    state of each job it was asked to keep."""
    from qoa_amd import StateVector

    handles = {n: StateVector(n) for n in sorted({n for n, _ in jobs})}
    try:
        for n, sv in handles.items():  # warm every handle's kernels once
            sv.run_gates(next(g for m, g in jobs if m == n)[:8])
            sv.synchronize()
        t0 = time.perf_counter()
        for n, g in jobs:
            sv = handles[n]
            sv.apply_reset_all()
            sv.run_gates(g)
            sv.synchronize()
        return time.perf_counter() - t0, handles
    except Exception:
        for sv in handles.values():
            sv.close()
        raise


def extra_c4(args) -> dict:
    """1024 ansatz jobs, all states resident: one launch per pass round (FP64,
    fused register passes) against the plain per-gate per-job FP64 loop on
    the same jobs; errors of sampled jobs against that loop."""
    from qoa_amd import StateVector
    from qoa_amd import workloads as W
    from qoa_amd.batch import Batch

    t0 = time.perf_counter()
    jobs = [W.ansatz_job(args.seed + j) for j in range(1024)]
    gen = time.perf_counter() - t0
    from qoa_amd._native import QOA_BATCH_AUTOTUNE, QOA_C64, load

    lib = load()

    def timed_runs(b, reps=3):
        b.run()
        b.synchronize()  # warm-up (and graph capture)
        times = []
        for _ in range(reps):
            b.run()
            times.append(b.synchronize())
        return min(times)

    pj = [(n, g) for n, _, g in jobs]
    t0 = time.perf_counter()
    fast = QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG
    b = Batch(pj, tile_qubits=12, flags=QOA_BATCH_AUTOTUNE | fast)
    plan_s = time.perf_counter() - t0
    info = b.info()
    ms = timed_runs(b)
    with Batch(pj, tile_qubits=12, flags=QOA_BATCH_AUTOTUNE) as bp:
        ms_plain = timed_runs(bp)
    tiles = {}
    for j, (n, _, _) in enumerate(jobs):
        tiles.setdefault(n, b.job_plan(j)[0])
    # the same batch with every kernel launched from the host (no graph)
    lib.qoa_set_option(b"batch_graphs", 0)
    try:
        with Batch(pj, tile_qubits=12, flags=QOA_BATCH_AUTOTUNE | fast) as bd:
            ms_direct = timed_runs(bd)
    finally:
        lib.qoa_set_option(b"batch_graphs", 1)
    with Batch(pj, tile_qubits=12, flags=fast) as b12:
        ms_12 = timed_runs(b12)
    amp_updates = float(sum(len(g) * (1 << n) for n, _, g in jobs))
    out = {"workload": "1024 hardware-efficient-ansatz circuits, n~U{8..24}, depth~U{10..60}, "
                       "FP64, all states resident", "jobs": 1024,
           "state_gib": info["total_amps"] * 16 / 2**30, "rounds": info["rounds"],
           "launches": "one graph replay per run (the rounds' kernels captured once)",
           "arithmetic": "FP64 register passes with factored H/RX/RY/RZ and merged diagonal "
                         "runs (QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG, FMA; within 1e-12, "
                         "sampled jobs below); ms_per_batch_unfactored: exact arithmetic",
           "ms_per_batch": ms, "jobs_per_s": 1024 / (ms * 1e-3),
           "ms_per_batch_unfactored": ms_plain,
           "amp_updates_per_s": amp_updates / (ms * 1e-3),
           "hbm_gbps": info["alg_bytes"] / (ms * 1e-3) / 1e9,
           "autotuned_tile_by_n": {str(k): v for k, v in sorted(tiles.items())},
           "ms_per_batch_fixed_tile_12": ms_12,
           "ms_per_batch_direct_launches": ms_direct,
           "gen_s": gen, "plan_and_tune_s": plan_s}
    # FP32-fused form: complex64 states, 5-qubit blocks on matrix cores
    try:
        t0 = time.perf_counter()
        with Batch(pj, precision=QOA_C64) as bf:
            fuse_s = time.perf_counter() - t0
            fi = bf.info()
            ms32 = timed_runs(bf, reps=2)
            nblk = sum(bf.job_plan(j)[1] for j in range(1024))
            sample = list(range(0, 1024, 64))
            worst_inf, worst_d = 0.0, 0.0
            for j in sample:
                d, inf = compare_host(bf.download(j), b.download(j))
                worst_d, worst_inf = max(worst_d, d), max(worst_inf, inf)
        out["fp32_blocks"] = {
            "method": "complex64 states, dense 5-qubit blocks (32x32, FP32 MFMA), round r = "
                      "block r of every job", "ms_per_batch": ms32,
            "jobs_per_s": 1024 / (ms32 * 1e-3), "blocks": nblk, "rounds": fi["rounds"],
            "hbm_gbps": fi["alg_bytes"] / (ms32 * 1e-3) / 1e9, "fuse_s": fuse_s,
            "vs_fp64_batch": {"jobs_compared": len(sample), "max_abs_diff": worst_d,
                              "one_minus_fidelity_max": worst_inf}}
    except Exception as e:  # noqa: BLE001
        out["fp32_blocks"] = {"error": str(e)}
    if not args.no_plain:
        log("c4: plain per-gate per-job FP64 loop over the same 1024 jobs ...")
        pj = [(n, g) for n, _, g in jobs]
        plain_s, handles = plain_c4_loop(pj)
        try:
            # errors: the batch's result of sampled jobs against the plain
            # loop's (run again for those jobs; the loop keeps one handle per n)
            sample = sorted(set(list(range(0, 1024, 64)) +
                                [max(range(1024), key=lambda j: jobs[j][0])]))
            worst_d, worst_inf = 0.0, 0.0
            for j in sample:
                n, _, g = jobs[j]
                sv = handles[n]
                sv.apply_reset_all()
                sv.run_gates(g)
                d, inf = compare_host(b.download(j), sv.download())
                worst_d, worst_inf = max(worst_d, d), max(worst_inf, inf)
        finally:
            for sv in handles.values():
                sv.close()
        out["plain_per_gate_loop"] = {
            "method": "qoa_run_gates per job, one job after another (one sweep kernel per "
                      "gate, FP64), all 1024 jobs", "s": plain_s, "jobs_per_s": 1024 / plain_s}
        out["speedup_vs_plain_loop"] = (1024 / (ms * 1e-3)) / (1024 / plain_s)
        if "jobs_per_s" in out.get("fp32_blocks", {}):
            out["fp32_blocks"]["speedup_vs_plain_loop"] = \
                out["fp32_blocks"]["jobs_per_s"] / (1024 / plain_s)
        out["accuracy_vs_plain_loop"] = {"jobs_compared": len(sample), "max_abs_diff": worst_d,
                                         "one_minus_fidelity_max": worst_inf}
    b.close()
    return out


def serving_stream(args, n_jobs: int, rate: float):
    from qoa_amd import workloads as W

    rng = np.random.default_rng(args.seed + int(rate))
    jobs = [W.serving_job(args.seed * 7 + j) for j in range(n_jobs)]
    gaps = rng.exponential(1.0 / rate, size=n_jobs)
    return jobs, np.cumsum(gaps)


def extra_c5_rate(args, rate: float, n_jobs: int) -> dict:
    from qoa_amd.server import Server

    jobs, offsets = serving_stream(args, n_jobs, rate)
    # factored forms (the pass scalars commute with the trajectory errors)
    with Server(budget_bytes=int(200 * 2**30), flags=QOA_FUSE_FACTOR) as srv:
        t_start = srv.now() + 0.2
        arrivals = t_start + offsets
        ids = [srv.submit(n, g, trajectories=64, p1=1e-3, p2=1e-2, noise_seed=args.seed + j,
                          arrival=float(arrivals[j]))
               for j, (n, _, g) in enumerate(jobs)]
        res = [srv.wait(i, timeout=600) for i in ids]
        stats = srv.stats()
    ok = [r for r in res if r.status == 0]
    lat = np.array([r.latency for r in ok])
    span = max(r.t_done for r in ok) - t_start
    window = float(offsets[-1])
    served = len(ok) / span
    return {"rate_jobs_per_s": rate, "jobs": n_jobs, "arrival_window_s": window,
            "completed": len(ok), "rejected": stats["rejected"], "jobs_per_s": served,
            "saturated": served < 0.9 * rate,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "steps": stats["steps"], "arena_peak_gib": stats["arena_peak"] / 2**30}


def plain_c5_loop(args) -> dict:
    """The plain per-gate per-job FP64 loop on the serving workload: a job's 64
    noisy trajectories one after another, each as qoa_run_trajectory (every
    gate's sweep, then the drawn Pauli sweeps). Its mean time per job is the
    expectation over the serving distribution P(n) ~ 2^-(n-8)/2, n = 8..26:
    one job of each size at the mean depth (35 layers), timed on 8
    trajectories (2 from 20 qubits up) and scaled to 64 (every trajectory
    costs the same)."""
    from qoa_amd import StateVector
    from qoa_amd import workloads as W

    w = {n: 2.0 ** (-(n - 8) / 2.0) for n in range(8, 27)}
    tot = sum(w.values())
    per_n = {}
    for n in range(8, 27):
        g = W.ansatz_gates(n, 35, W.SplitMix64(args.seed + n))
        traj = 8 if n < 20 else 2
        with StateVector(n) as sv:
            sv.run_trajectory(g[:1], 1, 0, 1e-3, 1e-2)  # warm
            sv.synchronize()
            t0 = time.perf_counter()
            for t in range(traj):
                sv.apply_reset_all()
                sv.run_trajectory(g, args.seed + n, t, 1e-3, 1e-2)
                sv.synchronize()
            per_n[n] = (time.perf_counter() - t0) / traj * 64
    mean_s = sum(w[n] / tot * per_n[n] for n in per_n)
    return {"method": "qoa_run_trajectory per trajectory, 64 per job, jobs one after another; "
                      "expected time per job over P(n), one 35-layer job per n timed",
            "s_per_job_by_n": per_n, "mean_s_per_job": mean_s, "jobs_per_s": 1.0 / mean_s}


def c5_accuracy(args) -> dict:
    """Served trajectories run the batched noisy kernel: for two serving jobs
    (64 trajectories each) the batched states against the plain per-gate
    trajectory loop in the same (planned) order, draws included."""
    from qoa_amd import StateVector
    from qoa_amd import workloads as W
    from qoa_amd.batch import Batch

    worst_d, worst_inf, cnt = 0.0, 0.0, 0
    j = 0
    picked = []
    while len(picked) < 2:
        n, _, g = W.serving_job(args.seed * 7 + j)
        if 10 <= n <= 16:
            picked.append((j, n, g))
        j += 1
    for j, n, g in picked:
        seed = args.seed + j
        with Batch([(n, g)], tile_qubits=12, trajectories=64, p1=1e-3, p2=1e-2,
                   noise_seed=seed) as b, StateVector(n) as sv:
            b.run()
            b.synchronize()
            order = b.order(0)
            for t in range(64):
                sv.apply_reset_all()
                sv.run_trajectory(g, seed, t, 1e-3, 1e-2, order=order)
                d, inf = compare_host(b.download(0, t), sv.download())
                worst_d, worst_inf = max(worst_d, d), max(worst_inf, inf)
                cnt += 1
    return {"trajectories_compared": cnt, "max_abs_diff": worst_d,
            "one_minus_fidelity_max": worst_inf}


def extra_a20(args) -> dict:
    """benches/benchmark.rs:91-252 as GPU units: per iteration the state is
    set to |0...0> on the device (the reference allocates it) and one gate is
    applied, target 0; n = 4..20. Also cpu-bench's loop (src/main.rs:
    1260-1311): H on qubit 0 through QuantumState::apply_h, i.e. the sweep and
    a normalize after every gate, at 20, 24 and 28 qubits."""
    from qoa_amd import QuantumState, StateVector

    units = {}
    for n in (4, 8, 12, 16, 20):
        reps = 200
        row = {}
        with StateVector(n) as sv:
            for name in A20_GATES:
                gpu_unit_gate(sv, name)
                sv.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    sv.apply_reset_all()
                    gpu_unit_gate(sv, name)
                sv.synchronize()
                row[name] = (time.perf_counter() - t0) / reps * 1e6
        units[f"{n}q_us_per_iter"] = row
    hb = {}
    for n in (20, 24, 28):
        with QuantumState(n) as qs:
            qs.apply_h(0)
            qs.synchronize()
            ops = 50
            t0 = time.perf_counter()
            for _ in range(ops):
                qs.apply_h(0)
            qs.synchronize()
            hb[f"{n}q_ops_per_s"] = ops / (time.perf_counter() - t0)
    units["cpu_bench_h_normalize"] = hb
    return units


def gpu_unit_gate(sv, name: str) -> None:
    if name == "hadamard":
        sv.apply_hadamard(0)
    elif name in ("rx", "ry", "rz"):
        getattr(sv, "apply_" + name)(0, math.pi / 2)
    elif name == "phaseshift":
        sv.apply_phaseshift(0, math.pi / 3)
    elif name == "reset":
        sv.apply_reset(0)
    elif name == "reset_all":
        sv.apply_reset_all()
    elif name in ("t", "s"):
        getattr(sv, "apply_" + name)(0)
    elif name == "cnot":
        sv.apply_cnot(0, 1)
    elif name == "swap":
        sv.apply_swap(0, 1)
    elif name == "cz":
        sv.apply_cz(0, 1)
    elif name == "cphase":
        sv.apply_controlled_phase_rotation(0, 1, math.pi / 4)
    elif name == "cswap":
        sv.apply_controlled_swap(2, 0, 1)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--method", default="fused-relabel",
                    choices=["fused", "fused-fma", "fused-merge", "fused-spec", "fused-factor",
                             "fused-relabel", "fused-lds", "per-gate"])
    ap.add_argument("--qubits", type=int, default=28)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--seed", type=int, default=20241203)
    ap.add_argument("--cpu-sample-gates", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--extras", default="",
                    help="comma list of extras to run (c1,c2,c3,c4,c5; default all)")
    ap.add_argument("--serve-rates", default="100,200,500,1000,2000,5000")
    ap.add_argument("--serve-jobs", type=int, default=1500,
                    help="jobs per arrival rate at most (at least 3 s of arrivals, >= 200)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain per-gate per-job loops of C4/C5")
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
    spec0 = specialize_info(lib)
    elapsed, kinds = run_circuit(sv, gates, args.method, args.steps, args.warmup, "c2")
    spec1 = specialize_info(lib)
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
    if spec1["compiles"] > spec0["compiles"] or args.method in ("fused-spec", "fused-factor",
                                                               "fused-relabel"):
        extras["c2_specialize"] = {
            "programs_compiled": spec1["compiles"] - spec0["compiles"],
            "compile_s": spec1["compile_seconds"] - spec0["compile_seconds"],
            "note": "hipRTC compile of the circuit's per-pass kernels, in the first (warm-up) "
                    "call and outside the timed steps; cached per process and on disk by "
                    "the generated source"}
    accuracy = None
    copy = None
    if wanted("c2"):
        # per-gate sweeps once on a second state: the reference every fused
        # method is compared against (bit-identical to the CPU oracle here)
        ref = StateVector(n)
        el, kk = run_circuit(ref, gates, "per-gate", 1, 0, "c2 per-gate")
        extras["c2_per-gate"] = {"ms_per_step": el * 1e3, "amp_updates_per_s": amp_updates / el,
                                 "kernels": kk}
        for other in ("fused", "fused-fma", "fused-merge", "fused-spec", "fused-factor",
                      "fused-relabel", "fused-lds"):
            el, kk = run_circuit(sv, gates, other, 1, 0, f"c2 {other}")
            d, f = sv.compare(ref)
            entry = {"ms_per_step": el * 1e3, "amp_updates_per_s": amp_updates / el,
                     "max_abs_diff_vs_per_gate": d, "one_minus_fidelity_vs_per_gate": 1.0 - f,
                     "kernels": kk}
            if other == args.method:
                accuracy = {"vs": "per-gate FP64 sweeps of the same circuit (bit-identical to "
                                  "the CPU oracle at 28 qubits)",
                            "max_abs_diff": d, "one_minus_fidelity": 1.0 - f}
            else:
                extras[f"c2_{other}"] = entry
        copy = copy_ceiling(ref)
        ref.close()
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
            nj = min(args.serve_jobs, max(200, int(3 * r)))
            try:
                serve.append(extra_c5_rate(args, r, nj))
                log(f"c5 serving: {serve[-1]}")
            except Exception as e:  # noqa: BLE001
                serve.append({"rate_jobs_per_s": r, "error": str(e)})
                log(f"c5 serving failed: {e}")
        c5 = {"workload": "Poisson arrivals of ansatz jobs, n in 8..26 with P(n) ~ "
                          "2^-(n-8)/2, depth 10..60, 64 noisy trajectories each (p1 1e-3, "
                          "p2 1e-2, operand channel), one sample per trajectory",
              "rates": serve}
        if not args.no_plain:
            try:
                plain = plain_c5_loop(args)
                best = max((x.get("jobs_per_s", 0.0) for x in serve), default=0.0)
                c5.update({"plain_per_gate_loop": plain,
                           "speedup_vs_plain_loop": best / plain["jobs_per_s"],
                           "accuracy_vs_plain_loop": c5_accuracy(args)})
                log(f"c5 plain loop: {plain['jobs_per_s']:.2f} jobs/s")
            except Exception as e:  # noqa: BLE001
                c5["plain_error"] = str(e)
                log(f"c5 plain loop failed: {e}")
        extras["c5_serving"] = c5
    if wanted("a20"):
        try:
            extras["a20_units"] = extra_a20(args)
        except Exception as e:  # noqa: BLE001
            extras["a20_units"] = {"error": str(e)}

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
                   "gates": len(gates), "seed": args.seed, "method": args.method,
                   "method_note": METHOD_NOTES.get(args.method, "")},
        "accuracy": accuracy,
        "roofline": roofline(dom, args, copy),
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


def roofline(dom, args, copy=None) -> dict:
    kind, k = dom
    per_launch = k["bytes"] / k["launches"]
    out = {"bound": "hbm", "achieved": k["gbps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": k["gbps"] / HBM_PEAK_GBS, "traffic": None, "kernel": kind,
           "avg_launch_us": k["avg_us"], "bytes_per_launch": per_launch}
    if copy:
        out["copy_ceiling"] = copy
        out["frac_of_copy"] = k["gbps"] / copy["gbps"]
        # whole-step view: passes per step and the step time if every pass
        # moved its bytes at the copy sweep's rate (fewer, fuller passes lower
        # the per-launch fraction while the step gets faster)
        per_step = k["launches"] / max(args.steps, 1)
        out["launches_per_step"] = per_step
        out["ms_per_step_at_copy_rate"] = per_step * per_launch / (copy["gbps"] * 1e9) * 1e3
    default = (args.qubits, args.depth, args.seed) == (28, 40, 20241203)
    key = {"fused": "fused_pass/28q-d40-ladders",
           "fused-merge": "fused_pass_merge/28q-d40-ladders",
           "fused-spec": "fused_pass_spec/28q-d40-ladders",
           "fused-factor": "fused_pass_factor/28q-d40-ladders",
           "fused-relabel": "fused_pass_relabel/28q-d40-ladders"}.get(args.method)
    prof = measured_profile(key) if (default and kind == "fused_pass" and key) else None
    if prof:
        out["traffic"] = prof["bytes_per_launch"]
        out["traffic_source"] = "profiles/traffic.json (rocprofv3 FETCH_SIZE/WRITE_SIZE, " \
                                "copy-calibrated)"
        # the pass is bound by instruction issue, not HBM: its FP64 rate and the
        # share of the pass's cycles each SIMD spends issuing VALU / SALU work
        flop = prof.get("fp64_flop", prof["fp64_valu_instr"] * 64)  # an FMA is two flops
        fl = flop / (k["avg_us"] * 1e-6) / 1e12
        out["fp64_valu"] = {"achieved_tflops": fl, "peak_tflops": FP64_TFLOPS,
                            "frac": fl / FP64_TFLOPS,
                            "fp64_share_of_valu_instr": prof["fp64_valu_instr"] / prof["valu_instr"]}
        cycles = k["avg_us"] * CLOCK_MHZ
        out["issue"] = {"valu_busy": prof["valu_instr"] / SIMDS * 4 / cycles,
                        "salu_busy": prof["salu_instr"] / SIMDS * 4 / cycles}
    return out


def copy_ceiling(sv, launches: int = 6) -> dict:
    """The practical HBM ceiling on this state: X on the top qubit reads and
    writes every amplitude once (a copy of 2^n x 16 B each way)."""
    n = sv.n
    sv.apply_x(n - 1)
    sv.synchronize()
    sv.timing_enable(True)
    for _ in range(launches):
        sv.apply_x(n - 1)
    sv.synchronize()
    perm = kernel_table(sv)["sweep_perm"]
    sv.timing_enable(False)
    return {"kernel": f"X sweep on qubit {n - 1} ({n} qubits, read + write 2^n x 16 B)",
            "gbps": perm["gbps"], "frac": perm["gbps"] / HBM_PEAK_GBS, "avg_us": perm["avg_us"]}


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
