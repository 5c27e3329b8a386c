"""Open-loop serving driver: Poisson arrivals into a kate_amd.Server.

Requests are built before the clock starts; each is submitted at its seeded
arrival time (non-blocking) and its latency is measured by the server from
submission to completion (queueing, planning, admission and device time
included).
"""
from __future__ import annotations

import math
import time
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import Circuit, Server
from . import workloads as W


def request_mix(count: int, seed: int, nmin: int = 12, nmax: int = 30, fuse: int = 0):
    """Config-C5 mix: n = round(exp(U[ln nmin, ln nmax])); half random
    layered circuits (depth U{10..40}), half Kate's QFT on a random basis
    input; 8 seeded Z-string masks each.  Returns [(circuit, masks, theta_phi)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        masks = [int(m) for m in W.zstring_masks(n, 8, seed * 7919 + i)]
        if i % 2 == 0:
            gates = W.layered(n, int(rng.integers(10, 41)), seed * 104729 + i)
            tp = None
        else:
            gates = W.qft(n)
            tp = W.basis_input(n, int(rng.integers(0, 1 << n)))
        c = Circuit(n, gates)
        if fuse:
            c.schedule().fuse(fuse)
        out.append((c, masks, tp))
    return out


def serve_poisson(requests, qps: float, seed: int, dtype: str = "c128", **server_kw) -> dict:
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, size=len(requests))
    arrivals = np.cumsum(gaps) - gaps[0]
    with Server(dtype=dtype, **server_kw) as srv:
        t0 = time.perf_counter()
        ids = []
        for (c, masks, tp), at in zip(requests, arrivals):
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(c, masks, tp))
        results = [srv.wait(i) for i in ids]
    lat = np.array([st["latency_ms"] for _, st in results])
    first = min(st["submit_s"] for _, st in results)
    last = max(st["done_s"] for _, st in results)
    return {
        "requests": len(requests), "offered_qps": qps, "dtype": dtype,
        "circuits_per_s": round(len(requests) / (last - first), 3),
        "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
        "mean_ms": round(float(lat.mean()), 2), "max_ms": round(float(lat.max()), 2),
        "qubits_served": int(sum(st["nqubits"] for _, st in results)),
        "results": [z for z, _ in results],
    }
