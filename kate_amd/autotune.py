"""Autotuner for batches (kate_batch_*): picks the GateMerger fusion width
(Kate's gateFusion qbitsizelimit, Circuit.hpp:99), the tile shape (tile
qubits and the low qubits every tile holds) and the persistent grid.

Every candidate is timed on the same seeded sample of the batch's circuits,
as a captured graph replayed ``reps`` times (best time kept).  The knobs are
searched one after the other (fusion width, tile qubits, low qubits,
workgroups per CU), each with the best values found so far.  A fusion width
whose fused programs are gate for gate the unfused ones is not timed: it
would run the same plan (on Haar brickwork Kate's cost model keeps every
2-qubit gate apart, so every width is inert there).  Timing is the only
criterion: every candidate computes the same circuits (fusion changes
rounding, not results).
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import Batch, Circuit


def build(work: Sequence[Tuple[int, list]], k: int) -> List[Circuit]:
    """Circuits of ``work`` ([(n, gates)]), scheduled and fused to width k
    (k = 0: no fusion)."""
    out = []
    for n, gates in work:
        c = Circuit(n, gates)
        if k:
            c.schedule().fuse(k)
        out.append(c)
    return out


def _signature(circuits: Sequence[Circuit]):
    return tuple(tuple(tuple(g.qubits) for g in c.gates()) for c in circuits)


def autotune_batch(work: Sequence[Tuple[int, list]], dtype: str = "c64", ks=(0, 3, 4, 5, 6), tiles=(11, 12, 13),
                   lows=(0, 3, 5), wgs=(4, 8, 16), sample: int = 256, reps: int = 2,
                   seed: int = 0) -> Tuple[Dict, List[Dict]]:
    """Returns (best candidate, table of every timed candidate); a candidate
    is {"k", "tile", "low", "wgs_per_cu", "ms"} with ms on the sample, and
    the table also lists inert fusion widths with "ms": None."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(work), size=min(sample, len(work)), replace=False)
    part = [work[int(i)] for i in idx]
    table: List[Dict] = []
    cache: Dict[Tuple, float] = {}
    built: Dict[int, List[Circuit]] = {}

    def time(k, tile, low, wpc):
        key = (k, tile, low, wpc)
        if key not in cache:
            with Batch(built[k], dtype=dtype, tile_qubits=tile, low_qubits=low, wgs_per_cu=wpc) as b:
                b.capture()
                b.run()
                cache[key] = min(b.run() for _ in range(reps))
            table.append({"k": k, "tile": tile, "low": low, "wgs_per_cu": wpc, "ms": cache[key]})
        return cache[key]

    best = {"k": 0, "tile": 12, "low": 0, "wgs_per_cu": 8}
    sigs = {}
    for k in ks:
        built[k] = build(part, k)
        sig = _signature(built[k])
        same = next((kk for kk, s in sigs.items() if s == sig), None)
        if same is not None:
            table.append({"k": k, "tile": None, "low": None, "wgs_per_cu": None, "ms": None,
                          "same_program_as_k": same})
            continue
        sigs[k] = sig
    live = sorted(sigs)
    best["k"] = min(live, key=lambda k: time(k, best["tile"], best["low"], best["wgs_per_cu"]))
    best["tile"] = min(tiles, key=lambda t: time(best["k"], t, best["low"], best["wgs_per_cu"]))
    best["low"] = min(lows, key=lambda lo: time(best["k"], best["tile"], lo, best["wgs_per_cu"]))
    best["wgs_per_cu"] = min(wgs, key=lambda w: time(best["k"], best["tile"], best["low"], w))
    best["ms"] = cache[(best["k"], best["tile"], best["low"], best["wgs_per_cu"])]
    return best, table
