"""Headline benchmark: one MI355X simulating Kate's single-GPU simulate call.

A step is one full simulate of the 30-qubit, depth-20 layered circuit (RX,
RZ on every qubit then the CNOT ladder, 1780 gates) from |0...0> in
complex128: initialise, run every pass, reduce Kate's proba_state.  The state
(16 GiB) and the program are resident in HBM before timing starts.

Prints one JSON line.  `value` is the FP64 circuit time in ms.  Beside it:
the per-pass roofline of the dominant kernel (k_pass, HIP events on the
simulator's stream), the complex64 + GateMerger(k=5) form of the same
circuit with its 1 - fidelity against the FP64 state, and the oracle (plain-C restatement of Kate's CPU simulator) timed
on this host's cores on a bounded sample.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import kate_amd as K  # noqa: E402
from kate_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per k_pass launch of the roofline set (the passes of the headline
# execute() that read and write every amplitude), from rocprofv3 FETCH_SIZE
# and WRITE_SIZE runs of `bench.py --passes-only execute`, each scaled by what
# the same counter reads on the k_copy_tile launches of that run (16 GiB read
# + 16 GiB written per launch: factors 2.000 and 1.000): 1.00002 x the
# algorithmic bytes (profiles/k_pass_f64_30q_layered_hbm_counters.md).  The
# ratio holds for the plan it was measured on; the bench reports it only when
# the current plan of the default workload is that plan.
TRAFFIC_RATIO_30Q_LAYERED = 1.00002
TRAFFIC_PLAN_30Q_LAYERED = "nq_int=30 tile_bits=12 low=3 passes=24 (gate 24, remap 0) micro=151 ops=199"
METRIC = "30q circuit ms + HBM % peak; served circuits/s + p99 at fidelity≥1-1e-6"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


FP64_VALU_PEAK_TFLOPS = 78.6


def plan_flops(circuit, dtype, n_int):
    """VALU flops of each pass of the device plan as executed over every
    tile, 2 per FMA-class instruction, per amplitude: 8 * 2^k for a dense
    k-qubit op (composed 1-qubit runs count once; a plain Kronecker op of m
    one-qubit gates counts m dense 1-qubit gates); 8 per bit plus 8 per
    diagonal table for a Kronecker op in Euler form, 4 per bit plus 8 in its
    scaled form; sparse ops (H, CNOT,
    CRk, Toffoli, phases) count 0.  Returns a list, one entry per pass."""
    os.environ["KATE_DESCRIBE_OPS"] = "1"
    try:
        text = circuit.plan(dtype)[2]
    finally:
        os.environ.pop("KATE_DESCRIBE_OPS", None)
    per_pass = []
    for ln in text.splitlines():
        if ln.startswith("pass "):
            per_pass.append(0.0)
            continue
        if " reg bits=" not in ln and " generic bits=" not in ln and " mfma" not in ln:
            continue
        head = ln.split(" tcols=")[0]
        f = 0.0
        for kind, code in re.findall(r" (dense|kron|ekron|fkron):(\d+)", head):
            m = int(code)
            k = bin(m & 15).count("1")
            if kind == "fkron":
                f += 4.0 * (1 << n_int) * (k + 2)
            elif kind == "ekron":
                f += 8.0 * (1 << n_int) * (k + 1 + (m >> 4))
            else:
                f += 8.0 * (1 << n_int) * (2 * k if kind == "kron" else (1 << k))
        if " generic bits=[" in ln or " mfma" in ln:
            k = len(ln.split("bits=[")[1].split("]")[0].split(","))
            f += 8.0 * (1 << n_int) * (1 << k)
        per_pass[-1] += f
    return per_pass


def pass_read_fraction(circuit, dtype):
    """Per pass of an execute() from |0...0>: the fraction of the state its
    loads read.  Pass 0 makes its tiles from the basis state and reads
    nothing; a later pass skips the amplitudes on the state bits still known
    to be zero (zero_in=k in the plan text: a fraction 2^-k).  Also returns
    the plan's summary line."""
    text = circuit.plan(dtype)[2]
    frac = []
    for ln in text.splitlines():
        if ln.startswith("pass ") or ln.startswith("remap "):
            m = re.search(r" zero_in=(\d+)", ln)
            frac.append(0.0 if not frac else (0.5 ** int(m.group(1)) if m else 1.0))
    return np.array(frac), text.splitlines()[0]


def mfma_steps(circuit, dtype, mfma):
    """Matrix-core micro-steps in the device plan of `circuit` (the plan
    text lists its steps by kind only with KATE_DESCRIBE_OPS set)."""
    os.environ["KATE_DESCRIBE_OPS"] = "1"
    try:
        text = circuit.plan(dtype, mfma=mfma)[2]
    finally:
        os.environ.pop("KATE_DESCRIBE_OPS", None)
    return sum(1 for ln in text.splitlines() if re.match(r"\s+step \d+ mfma", ln))


def time_executes(sim, steps, warmup):
    for _ in range(warmup):
        sim.execute()
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.execute()
    sim.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def cpu_baseline(n, depth, budget_s=15.0):
    """The oracle (plain-C restatement of Kate's CPU simulator: 14-qubit
    group staging, one thread per group range) on the full-size state of the
    headline circuit, compiled as Kate's own driver compiles it
    (gateScheduling, gateGrouping(12), compileDefault(0); main.cpp:30-34).
    Groups run one at a time from |0...0> until `budget_s` of CPU work; each
    group is one staged pass over the whole state, so the circuit time is
    the measured time per group times the group count."""
    from oracle import oracle as O

    threads = min(16, len(os.sched_getaffinity(0)))
    gates = W.layered(n, depth, 1)
    c = K.Circuit(n, gates).schedule().group(12).compile_default(0)
    ip, fi = c.permutations()
    prog = O.Compiled(n, c.gates(), c.groups(), [(1, [])] * len(c.groups()), ip or list(range(n)),
                      fi or list(range(n)))
    st = O.init_basis(n)
    done, sec = 0, 0.0
    while done < len(prog.groups) and sec < budget_s:
        t0 = time.perf_counter()
        O.run_groups(prog, st, done, done + 1, nthreads=threads)
        sec += time.perf_counter() - t0
        done += 1
    del st
    total = sec * 1e3 * len(prog.groups) / done
    return {"value": round(total, 1), "unit": "ms", "cores": threads, "kind": "port",
            "sample": f"{n}-qubit depth-{depth} layered circuit (complex128, {(1 << n) * 16 / 2 ** 30:g} GiB state) compiled as Kate's "
                      f"driver does (schedule, group(12), compileDefault(0)): {len(prog.groups)} groups; the first "
                      f"{done} ran in {sec:.2f} s on {threads} threads and the value is that time per group times "
                      f"{len(prog.groups)}"}


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
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-qulacs", action="store_true")
    ap.add_argument("--no-dense5", action="store_true")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--serve-qps", type=float, default=16.0)
    ap.add_argument("--serve-qps-high", type=float, default=1000.0)
    ap.add_argument("--passes-only", choices=("execute", "dense", "mfma"), default=None,
                    help="for profilers: one execute() of the headline circuit (execute), or one run of every "
                         "pass over every tile (dense), after one untimed warm-up execute(); mfma: one execute() "
                         "of the 28-qubit layered circuit, complex64, GateMerger(5)-fused, on FP32 matrix-core "
                         "windows; prints the pass times")
    ap.add_argument("--dtype", choices=("c128", "c64"), default="c128", help="state type of --passes-only")
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
    if args.passes_only == "mfma":
        n, depth = 28, 20
        c = K.Circuit(n, W.layered(n, depth, 2)).schedule().fuse(5)
        with K.Simulator(c, dtype="c64", mfma=2) as sim:
            sim.execute()
            _, ms, tiles = sim.execute_timed()
            sim.synchronize()
        print(json.dumps({"mode": "mfma", "dtype": "c64", "qubits": n, "depth": depth,
                          "mfma_steps": mfma_steps(c, "c64", 2), "passes": len(ms),
                          "sum_ms": round(float(ms.sum()), 3), "pass_ms": [round(float(x), 3) for x in ms]}),
              flush=True)
        return
    if args.passes_only:
        with K.Simulator(K.Circuit(n, gates), dtype=args.dtype, support=args.passes_only == "execute") as sim:
            sim.execute()
            copy_ms = None
            if args.passes_only == "execute":
                _, ms, tiles = sim.execute_timed()
                # the same passes' tiles copied with no arithmetic and each
                # pass's own addressing (k_copy_tile in place, k_scatter for
                # out-of-place passes): the HBM time each pattern allows
                copy_ms = sim.copy_ceiling()
            else:
                sim.init_basis(0)
                ms, _ = sim.run_timed()
            sim.synchronize()
        print(json.dumps({"mode": args.passes_only, "dtype": args.dtype, "qubits": n, "depth": depth,
                          "passes": len(ms), "sum_ms": round(float(ms.sum()), 3),
                          "pass_ms": [round(float(x), 3) for x in ms],
                          "copy_ms": None if copy_ms is None else [round(float(x), 3) for x in copy_ms]}),
              flush=True)
        return

    # ---- FP64 plain path (headline) ----
    # execute() runs every pass of the circuit and leaves the finished state
    # in HBM, as Kate's execute() does; from |0...0> its passes follow the
    # state's support (tiles that are still exactly zero are not launched).
    c = K.Circuit(n, gates)
    sim = K.Simulator(c, dtype="c128")
    ms64 = time_executes(sim, args.steps, args.warmup)
    t = sim.timings()
    # the same execute() with an event after every pass: the roofline of
    # k_pass is quoted on its passes that read and write every amplitude
    _, pass_ms, tiles = sim.execute_timed()
    sim.synchronize()
    state_bytes = float(1 << max(n, 12)) * 16
    bpp = 2.0 * state_bytes
    read_frac, plan_line = pass_read_fraction(c, "c128")
    if len(read_frac) != len(pass_ms):
        raise RuntimeError(f"plan has {len(read_frac)} passes, execute_timed() {len(pass_ms)}")
    full = (tiles == tiles.max()) & (read_frac == 1.0)
    mean_pass = float(pass_ms[full].mean())
    achieved = bpp / (mean_pass * 1e-3) / 1e9
    # algorithmic bytes of the whole execute(): each pass reads its non-zero
    # input and writes every tile it launches
    circ_bytes = float(np.sum(state_bytes * (read_frac + tiles / tiles.max())))
    circuit_gbs = circ_bytes / (float(pass_ms.sum()) * 1e-3) / 1e9
    traffic = round(bpp * TRAFFIC_RATIO_30Q_LAYERED) if plan_line == TRAFFIC_PLAN_30Q_LAYERED else None
    log(f"fp64: {ms64:.2f} ms/circuit, {len(pass_ms)} passes ({int(full.sum())} read and write every "
        f"amplitude), mean of those {mean_pass:.3f} ms = {achieved:.0f} GB/s")
    # the HBM time each pass's access pattern allows: a plain in-place copy
    # of the same tiles (no arithmetic, no LDS), timed the same way
    ceil_ms = sim.copy_ceiling()
    ceil_full = float(ceil_ms[full].mean())
    log(f"fp64 copy ceiling of the same tiles: mean {ceil_full:.3f} ms = {bpp / (ceil_full * 1e-3) / 1e9:.0f} GB/s; "
        f"the full passes reach {ceil_full / mean_pass:.3f} of it")
    flops = plan_flops(c, "c128", max(n, 12))
    full_flops = float(sum(f for f, on in zip(flops, full) if on))
    tflops = full_flops / (float(pass_ms[full].sum()) * 1e-3) / 1e12
    log(f"fp64 VALU: {sum(flops) / 1e12:.2f} TFLOP planned, {tflops:.1f} TFLOP/s over the full passes")

    # beside it, not the headline: every tile of every pass (support skipping
    # off), and execute() leaving its trailing permutation passes for the
    # next read of the state (deferral on)
    with K.Simulator(K.Circuit(n, gates), dtype="c128", support=False) as dense:
        ms_dense = time_executes(dense, args.steps, args.warmup)
        _, dense_ms, _ = dense.execute_timed()
    dense_mean = float(dense_ms.mean())
    log(f"fp64 dense (support skipping off): {ms_dense:.2f} ms/circuit, mean pass {dense_mean:.3f} ms")
    # (the scatter store folds those passes into the last gate pass; the
    # deferral is measured on the plan that keeps them as passes, which is
    # also in place: relabelled passes need the same second buffer)
    with K.Simulator(K.Circuit(n, gates), dtype="c128", defer=True, scatter=False) as deferred:
        ms_deferred = time_executes(deferred, args.steps, args.warmup)
    log(f"fp64 in place with the trailing permutation passes kept as passes and deferred: "
        f"{ms_deferred:.2f} ms/circuit")

    out = {
        "metric": METRIC, "value": round(ms64, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms64, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n}-qubit depth-{depth} layered circuit (RX, RZ on every qubit, CNOT ladder), "
                               f"complex128, from |0...0>, every pass run inside execute()", "qubits": n,
                   "depth": depth, "gates": len(gates), "state_gib": (1 << n) * 16 / 2 ** 30},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "k_pass<double,12>", "algorithmic_bytes_per_launch": bpp,
                     "mean_launch_ms": round(mean_pass, 4), "launches": int(full.sum()),
                     "run": "the headline execute(): its passes that read and write every amplitude (the "
                            "first passes skip the tiles and amplitudes that are still zero)",
                     "copy_ceiling": {"mean_launch_ms": round(ceil_full, 4),
                                      "achieved": round(bpp / (ceil_full * 1e-3) / 1e9, 1),
                                      "frac_of_ceiling": round(ceil_full / mean_pass, 4),
                                      "what": "the same passes' tiles copied with each pass's own addressing and "
                                              "no arithmetic (kate_sim_copy_ceiling): the HBM rate their access "
                                              "pattern allows"},
                     "compute": {"bound": "fp64 valu", "achieved": round(tflops, 2),
                                 "peak": FP64_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                                 "frac": round(tflops / FP64_VALU_PEAK_TFLOPS, 4),
                                 "valu_flops": full_flops}},
        "passes": int(len(pass_ms)), "pass_ms": [round(float(x), 3) for x in pass_ms],
        "pass_tiles": [int(x) for x in tiles],
        "init_ms": round(t["init_ms"], 3), "compute_ms": round(t["compute_ms"], 3),
        "measure_ms": round(t["measure_ms"], 3),
        "circuit_hbm_frac": round(circuit_gbs / HBM_PEAK_GBS, 4),
        "fp64_dense_ms": round(ms_dense, 3), "fp64_dense_mean_pass_ms": round(dense_mean, 4),
        "fp64_deferred_ms": round(ms_deferred, 3),
    }

    # ---- complex64 + GateMerger(k=5) ----
    if not args.no_fp32:
        for label, circ in (("fp32_fused", K.Circuit(n, gates).schedule().fuse(5)),
                            ("fp32_plain", K.Circuit(n, gates))):
            with K.Simulator(circ, dtype="c64") as s32:
                ms32 = time_executes(s32, args.steps, args.warmup)
                t32 = s32.timings()
                # 1 - F against the FP64 state, reduced on the device in FP64.
                infid = 1.0 - sim.fidelity(s32)
            out[label] = {"ms": round(ms32, 3), "speedup_vs_fp64": round(ms64 / ms32, 3),
                          "passes": t32["passes"], "gates": t32["gates"], "one_minus_fidelity": infid}
            log(f"{label}: {ms32:.2f} ms/circuit ({ms64 / ms32:.2f}x), 1-F = {infid:.3e}")
    sim.close()

    if not args.no_c2:
        out["c2_28q"] = c2()
        log(f"c2: {out['c2_28q']}")
    if not args.no_c3:
        out["c3_brickwork"] = c3()
        log(f"c3: {out['c3_brickwork']}")
    if not args.no_dense5:
        out["dense5_28q"] = dense5()
        log(f"dense5_28q: {out['dense5_28q']}")
    if not args.no_qulacs:
        out["qulacs25"] = qulacs()
        log(f"qulacs25: {out['qulacs25']}")
    if not args.no_batch:
        out["batched"] = batched(args.batch_count)
        log(f"batched: {out['batched']}")
    if not args.no_serve:
        out["serving"] = serving(args.serve_requests, args.serve_qps, args.serve_qps_high)
        log(f"serving: {out['serving']}")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, depth)
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


def pass_rates(sim, state_bytes):
    """Per-pass HBM GB/s (2 * state bytes per pass over its event time),
    split into gate passes and remap passes."""
    sim.init_basis(0)
    ms, remap = sim.run_timed()
    sim.synchronize()
    out = {}
    for label, sel in (("gate_passes", ~remap), ("remap_passes", remap)):
        t = ms[sel]
        if len(t):
            gbs = 2.0 * state_bytes / (t * 1e-3) / 1e9
            out[label] = {"count": int(len(t)), "mean_ms": round(float(t.mean()), 4),
                          "mean_GBs": round(float(2.0 * state_bytes * len(t) / (t.sum() * 1e-3) / 1e9), 1),
                          "min_GBs": round(float(gbs.min()), 1), "max_GBs": round(float(gbs.max()), 1)}
    return out


def c2(steps=3):
    """28-qubit depth-20 layered circuit: FP64 plain beside complex64
    (plain, and GateMerger(k=5)-fused on the VALU register path, on FP32
    matrix-core windows and on split-bf16 matrix-core windows), with
    max-abs amplitude error and 1 - fidelity against the FP64 state and the
    number of matrix-core steps each plan holds."""
    n, depth = 28, 20
    gates = W.layered(n, depth, 2)
    res = {"qubits": n, "depth": depth, "gates": len(gates)}
    with K.Simulator(K.Circuit(n, gates), dtype="c128") as ref:
        res["fp64_ms"] = round(time_executes(ref, steps, 1), 3)
        res["fp64_passes"] = ref.timings()["passes"]
        legs = (("fp32_plain", False, 0), ("fp32_fused_valu", True, 0), ("fp32_fused_mfma_fp32", True, 2),
                ("fp32_fused_mfma_bf16x3", True, 3))
        for label, fuse, mfma in legs:
            c = K.Circuit(n, gates).schedule().fuse(5) if fuse else K.Circuit(n, gates)
            steps_mfma = mfma_steps(c, "c64", mfma)
            with K.Simulator(c, dtype="c64", mfma=mfma) as s32:
                ms = time_executes(s32, steps, 1)
                _, na, nb, mx = ref.compare(s32)
                res[label] = {"ms": round(ms, 3), "speedup_vs_fp64": round(res["fp64_ms"] / ms, 3),
                              "max_abs_err": mx, "one_minus_fidelity": 1.0 - ref.fidelity(s32),
                              "norm_fp32": nb, "mfma_steps": steps_mfma, "passes": s32.timings()["passes"]}
    return res


def c3(depth=24):
    """Haar brickwork, depth 24, long-range (q, q+16 mod n) pairs every 4th
    layer: 31 qubits complex128 and 32 qubits complex64 (32 GiB each), per-pass
    HBM GB/s; the complex64 circuit also on matrix-core windows (FP32 and
    split-bf16 products).  The same circuits compiled for a 3-qubit global
    block (Kate's allocate(3) swap plan) show the in-HBM remap passes that
    replace its multi-GPU qubit swaps."""
    res = {}
    for n, dtype, amp in ((31, "c128", 16), (32, "c64", 8)):
        gates = W.brickwork(n, depth, 5)
        state = (1 << n) * amp
        entry = {"qubits": n, "dtype": dtype, "gates": len(gates), "state_gib": state / 2 ** 30}
        with K.Simulator(K.Circuit(n, gates), dtype=dtype) as sim:
            entry["ms"] = round(time_executes(sim, 2, 1), 3)
            entry["passes"] = pass_rates(sim, state)
            if dtype == "c64":
                # the same circuit on 5-qubit matrix-core windows (FP32
                # products, and split-bf16 products), against the VALU state
                for label, mfma in (("mfma_fp32", 2), ("mfma_bf16x3", 3)):
                    c = K.Circuit(n, gates)
                    with K.Simulator(c, dtype=dtype, mfma=mfma) as sm:
                        ms = time_executes(sm, 2, 1)
                        mx = sim.compare(sm)[3]
                        entry[label] = {"ms": round(ms, 3), "speedup_vs_valu": round(entry["ms"] / ms, 3),
                                        "mfma_steps": mfma_steps(c, dtype, mfma), "passes": sm.timings()["passes"],
                                        "max_abs_err_vs_valu": mx,
                                        "one_minus_fidelity_vs_valu": 1.0 - sim.fidelity(sm)}
        with K.Simulator(K.Circuit(n, gates).schedule().group(8).allocate(3), dtype=dtype) as sim:
            entry["swap_plan_ms"] = round(time_executes(sim, 1, 1), 3)
            entry["swap_plan_passes"] = pass_rates(sim, state)
        res[f"{n}q_{dtype}"] = entry
    return res


def qulacs(steps=5):
    """Kate's QulacsBench(25) (1025 gates), FP64, compiled as its main
    program does (schedule, group(12), compile_default(0)); Kate publishes
    1035.41 ms for it on one MI200."""
    n = 25
    gates = W.qulacs_bench(n)
    c = K.Circuit(n, gates).schedule().group(12).compile_default(0)
    with K.Simulator(c, dtype="c128") as sim:
        ms = time_executes(sim, steps, 2)
        t = sim.timings()
    return {"qubits": n, "gates": len(gates), "ms": round(ms, 3), "passes": t["passes"],
            "kate_mi200_ms": 1035.41, "speedup_vs_kate_mi200": round(1035.41 / ms, 1)}


def dense5(steps=3):
    """Kernel family (b) where it pays: 28 qubits, 60 Haar-random dense
    5-qubit gates on random qubits.  FP64 (register-blocked VALU steps),
    complex64 as planned by default (the 5-qubit gates on the FP32 matrix
    cores) and complex64 with them on the VALU; 1-F against the FP64 state."""
    n = 28
    rng = np.random.default_rng(75)
    gates = [W.Gate(W.DENSE, tuple(int(x) for x in rng.choice(n, size=5, replace=False)), 0,
                    W.haar_unitary(rng, 32)) for _ in range(60)]
    res = {"qubits": n, "gates": len(gates)}
    with K.Simulator(K.Circuit(n, gates), dtype="c128") as ref:
        res["fp64_ms"] = round(time_executes(ref, steps, 1), 3)
        for label, m5 in (("fp32_matrix_cores", True), ("fp32_valu", False)):
            with K.Simulator(K.Circuit(n, gates), dtype="c64", matrix5=m5) as s32:
                ms = time_executes(s32, steps, 1)
                os.environ["KATE_DESCRIBE_OPS"] = "1"
                try:
                    text = K.Circuit(n, gates).plan("c64", matrix5=m5)[2]
                finally:
                    os.environ.pop("KATE_DESCRIBE_OPS", None)
                res[label] = {"ms": round(ms, 3), "speedup_vs_fp64": round(res["fp64_ms"] / ms, 3),
                              "one_minus_fidelity": float(1.0 - s32.fidelity(ref)),
                              "mfma_steps": sum(1 for ln in text.splitlines() if re.match(r"\s+step \d+ mfma", ln))}
    return res


def batched(count, check=64):
    """Config C4: `count` Haar brickwork circuits (n log-uniform in [10, 20],
    depth U{8..32}), complex64, one persistent launch per run under a captured
    graph; fusion width, tile shape and persistent grid from the autotuner on
    a 256-circuit sample.  <Z_q> of `check` seeded circuits against the FP64
    oracle, and of every circuit against the same batch run in FP64."""
    from kate_amd.autotune import autotune_batch, build
    from oracle import oracle as O

    t0 = time.perf_counter()
    work = W.batch_workload(count, seed=4)
    best, table = autotune_batch(work, dtype="c64", sample=256)
    circuits = build(work, best["k"])
    with K.Batch(circuits, dtype="c64", tile_qubits=best["tile"], low_qubits=best["low"],
                 wgs_per_cu=best["wgs_per_cu"]) as b:
        b.capture()
        b.run()
        ms = min(b.run() for _ in range(3))
        info = b.info()
        z32, _ = b.expect_z()
    with K.Batch([K.Circuit(n, g) for n, g in work], dtype="c128") as b64:
        b64.run()
        z64, _ = b64.expect_z()
    zerr = max(float(np.abs(np.asarray(a, dtype=np.float64) - np.asarray(r)).max()) for a, r in zip(z32, z64))
    rng = np.random.default_rng(7)
    oerr = 0.0
    for i in rng.choice(len(work), size=min(check, len(work)), replace=False):
        n, g = work[int(i)]
        ref = O.expect_z(n, O.simulate(n, g), [1 << q for q in range(n)])
        oerr = max(oerr, float(np.abs(z32[int(i)] - ref).max()))
    return {"circuits": count, "ms": round(ms, 3), "circuits_per_s": round(count / (ms * 1e-3), 1),
            "max_z_err_vs_fp64_batch": zerr, "max_z_err_vs_oracle": oerr, "oracle_checked": min(check, count),
            "autotuned": {k: v for k, v in best.items() if k != "ms"}, "tasks": info["tasks"],
            "persistent_grid": info["grid"], "autotune_table": table,
            "setup_s": round(time.perf_counter() - t0, 1)}


def serving(nreq, qps, qps_high, fid_sample=24):
    """Config C5: Poisson arrivals of layered / QFT circuits, n log-uniform
    in [12, 30], 8 Z strings each; FP64 plain and complex64 + GateMerger(k=5),
    at a light load (latency) and at a load above capacity (circuits/s the
    continuous-batching loop sustains).  Accuracy of the FP32 server per
    request: max |<Z_S>| difference against the FP64 server, and, on a
    seeded sample of the requests, 1 - fidelity of the complex64 state
    against the complex128 one (the same fused circuit and input, reduced on
    the device)."""
    from kate_amd.serving import request_mix, serve_poisson

    res, zs = {}, {}
    for dtype, fuse in (("c128", 0), ("c64", 5)):
        label = "fp64_plain" if dtype == "c128" else "fp32_fused"
        reqs = request_mix(nreq, seed=11, fuse=fuse)
        r = serve_poisson(reqs, qps, seed=12, dtype=dtype, streams=4, max_active=32)
        zs[dtype] = r.pop("results")
        res[label] = r
        r = serve_poisson(reqs, qps_high, seed=13, dtype=dtype, streams=4, max_active=32)
        r.pop("results")
        res[label + "_saturated"] = r
        log(f"serving {label}: {res[label]}; saturated: {r}")
    per_req = np.array([float(np.abs(np.asarray(a) - np.asarray(b)).max()) for a, b in zip(zs["c64"], zs["c128"])])
    res["fp32_fused"]["z_err_vs_fp64"] = {"max": float(per_req.max()), "p99": float(np.percentile(per_req, 99)),
                                         "p50": float(np.percentile(per_req, 50))}
    reqs = request_mix(nreq, seed=11, fuse=5)
    rng = np.random.default_rng(17)
    infid = []
    for i in sorted(rng.choice(len(reqs), size=min(fid_sample, len(reqs)), replace=False)):
        c, _, tp = reqs[int(i)]
        with K.Simulator(c, dtype="c128") as a, K.Simulator(c, dtype="c64") as b:
            a.execute(tp)
            b.execute(tp)
            infid.append(1.0 - a.fidelity(b))
    infid = np.array(infid)
    res["fp32_fused"]["one_minus_fidelity"] = {"sample": int(len(infid)), "max": float(infid.max()),
                                               "p50": float(np.percentile(infid, 50))}
    return res


if __name__ == "__main__":
    main()
