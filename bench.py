"""Headline benchmark: one MI355X simulating Kate's single-GPU simulate call.

A step is one full simulate of the 30-qubit, depth-20 layered circuit (RX,
RZ on every qubit then the CNOT ladder, 1780 gates) from |0...0> in
complex128: initialise, run every pass, reduce Kate's proba_state.  The state
(16 GiB) and the program are resident in HBM before timing starts.

Prints one JSON line.  `value` is the FP64 circuit time in ms.  Beside it:
the per-pass roofline of the dominant kernel (k_pass, HIP events on the
simulator's stream), the complex64 + GateMerger(k=5) form of the same
circuit, and the oracle (plain-C restatement of Kate's CPU simulator) timed
on this host's cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import kate_amd as K  # noqa: E402
from kate_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "30q circuit ms + HBM % peak; served circuits/s + p99 at fidelity≥1-1e-6"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_executes(sim, steps, warmup):
    for _ in range(warmup):
        sim.execute()
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.execute()
    sim.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def cpu_baseline(n_sample, depth, n_full, gates_full):
    """The oracle (restated Kate CPU simulator, 14-qubit group staging, one
    thread per group range) on a bounded sample, extrapolated linearly in
    gate-amplitude updates to the full circuit."""
    from oracle import oracle as O

    threads = min(16, len(os.sched_getaffinity(0)))
    gates = W.layered(n_sample, depth, 1)
    t0 = time.perf_counter()
    O.simulate(n_sample, gates, nthreads=threads)
    sec = time.perf_counter() - t0
    work_sample = len(gates) * (1 << n_sample)
    work_full = gates_full * (1 << n_full)
    return {"value": round(sec * 1e3 * work_full / work_sample, 1), "unit": "ms", "cores": threads, "kind": "port",
            "sample": f"{n_sample}-qubit depth-{depth} layered circuit ({len(gates)} gates) took {sec:.2f} s; "
                      f"scaled by gate-amplitude updates to the {n_full}-qubit circuit"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-qps", type=float, default=4.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the simulate path it measures is Kate's single-GPU call "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    if K.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    n, depth = args.qubits, args.depth
    gates = W.layered(n, depth, 1)
    log(f"device: {K.device_info().splitlines()[0]}")

    # ---- FP64 plain path (headline) ----
    c = K.Circuit(n, gates)
    with K.Simulator(c, dtype="c128") as sim:
        ms64 = time_executes(sim, args.steps, args.warmup)
        t = sim.timings()
        sim.init_basis(0)
        pass_ms, remap = sim.run_timed()
        sim.synchronize()
    bpp = 2.0 * (1 << max(n, 12)) * 16
    gate_ms = pass_ms[~remap]
    mean_pass = float(gate_ms.mean())
    achieved = bpp / (mean_pass * 1e-3) / 1e9
    circuit_gbs = bpp * len(pass_ms) / (t["compute_ms"] * 1e-3) / 1e9
    log(f"fp64: {ms64:.2f} ms/circuit, {len(pass_ms)} passes, mean pass {mean_pass:.3f} ms, "
        f"{achieved:.0f} GB/s per pass")

    out = {
        "metric": METRIC, "value": round(ms64, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms64, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n}-qubit depth-{depth} layered circuit (RX, RZ on every qubit, CNOT ladder), "
                               f"complex128, from |0...0>", "qubits": n, "depth": depth, "gates": len(gates),
                   "state_gib": (1 << n) * 16 / 2 ** 30},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_pass<double,12>", "algorithmic_bytes_per_launch": bpp,
                     "mean_launch_ms": round(mean_pass, 4), "launches": int(len(gate_ms))},
        "passes": int(len(pass_ms)), "init_ms": round(t["init_ms"], 3), "compute_ms": round(t["compute_ms"], 3),
        "measure_ms": round(t["measure_ms"], 3), "circuit_hbm_frac": round(circuit_gbs / HBM_PEAK_GBS, 4),
    }

    # ---- complex64 + GateMerger(k=5) ----
    if not args.no_fp32:
        c32 = K.Circuit(n, gates).schedule().fuse(5)
        with K.Simulator(c32, dtype="c64") as sim:
            ms32 = time_executes(sim, args.steps, args.warmup)
            t32 = sim.timings()
        out["fp32_fused"] = {"ms": round(ms32, 3), "speedup_vs_fp64": round(ms64 / ms32, 3),
                             "passes": t32["passes"], "gates_after_fusion": t32["gates"]}
        log(f"fp32 fused: {ms32:.2f} ms/circuit ({ms64 / ms32:.2f}x)")

    if not args.no_batch:
        out["batched"] = batched(args.batch_count)
        log(f"batched: {out['batched']}")
    if not args.no_serve:
        out["serving"] = serving(args.serve_requests, args.serve_qps)
        log(f"serving: {out['serving']}")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(min(n, 22), depth, n, len(gates))
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


def batched(count):
    """Config C4: `count` Haar brickwork circuits (n log-uniform in [10, 20],
    depth U{8..32}), complex64, one persistent launch per run under a captured
    graph; fusion width and tile shape from the autotuner on a 256-circuit
    sample."""
    from kate_amd.autotune import autotune_batch, build

    t0 = time.perf_counter()
    work = W.batch_workload(count, seed=4)
    (k, tile, sample_ms), table = autotune_batch(work, dtype="c64", ks=(3, 4, 5, 6), tiles=(11, 12, 13), sample=256)
    circuits = build(work, k)
    with K.Batch(circuits, dtype="c64", tile_qubits=tile) as b:
        b.capture()
        b.run()
        ms = min(b.run() for _ in range(3))
        info = b.info()
    return {"circuits": count, "ms": round(ms, 3), "circuits_per_s": round(count / (ms * 1e-3), 1),
            "fusion_k": k, "tile_qubits": tile, "tasks": info["tasks"], "persistent_grid": info["grid"],
            "autotune_table_ms": [[kk, tt, round(mm, 3)] for kk, tt, mm in table],
            "setup_s": round(time.perf_counter() - t0, 1)}


def serving(nreq, qps):
    """Config C5: Poisson arrivals of layered / QFT circuits, n log-uniform
    in [12, 30], 8 Z strings each; FP64 plain and complex64 + GateMerger(k=5)."""
    from kate_amd.serving import request_mix, serve_poisson

    res = {}
    for dtype, fuse in (("c128", 0), ("c64", 5)):
        reqs = request_mix(nreq, seed=11, fuse=fuse)
        r = serve_poisson(reqs, qps, seed=12, dtype=dtype, streams=4, max_active=32)
        r.pop("results")
        res["fp64_plain" if dtype == "c128" else "fp32_fused"] = r
    return res


if __name__ == "__main__":
    main()
