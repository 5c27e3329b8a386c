"""Autotuner for batches (kate_batch_*): picks the GateMerger fusion width
(Kate's gateFusion qbitsizelimit, Circuit.hpp:99) and the tile shape.

Each candidate (k, tile) is timed on the same seeded sample of the batch's
circuits, as a captured graph replayed ``reps`` times (best time kept); the
fastest candidate wins.  Timing is the only criterion: every candidate
computes the same circuits (fusion changes rounding, not results).
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

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


def autotune_batch(work: Sequence[Tuple[int, list]], dtype: str = "c64", ks=(3, 4, 5, 6), tiles=(11, 12, 13),
                   sample: int = 256, reps: int = 2, seed: int = 0):
    """Returns ((k, tile, ms_on_sample), table of every candidate)."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(work), size=min(sample, len(work)), replace=False)
    part = [work[int(i)] for i in idx]
    table = []
    for k in ks:
        circuits = build(part, k)
        for tile in tiles:
            with Batch(circuits, dtype=dtype, tile_qubits=tile) as b:
                b.capture()
                b.run()
                ms = min(b.run() for _ in range(reps))
            table.append((k, tile, ms))
    best = min(table, key=lambda x: x[2])
    return best, table
