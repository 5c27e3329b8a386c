"""Writes the golden fixtures of tests/golden/ from Kate's own CPU build.

Run in a checkout where oracle/_ref/libkateref.so was built from the
reference sources (``make -C oracle ref``):

    python tests/golden/gen_golden.py

Each fixture holds the input gates (matrices included, so no generator has to
be bit-reproducible) and Kate's outputs: the compiled program after the named
passes, and/or the final amplitudes and proba_state of Kate's CPUSimulator.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from fixture_io import pack_gates, pack_program  # noqa: E402
from kate_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

PIPELINES = {
    "raw": [],
    "sched": ["schedule"],
    "fuse5": ["fuse5"],
    "sched_fuse5": ["schedule", "fuse5"],
    "sched_fuse3_group6": ["schedule", "fuse3", "group6"],
    "sched_group5_alloc2": ["schedule", "group5", "allocate2"],
    "sched_fuse4_group8_cd0": ["schedule", "fuse4", "group8", "cd0"],
    "alloc0": ["allocate0"],
    "group4_cd2": ["group4", "cd2"],
}


def apply(ref, steps):
    for s in steps:
        if s == "schedule":
            ref.schedule()
        elif s.startswith("fuse"):
            ref.fuse(int(s[4:]), 1e-6)
        elif s.startswith("group"):
            ref.group(int(s[5:]))
        elif s.startswith("allocate"):
            ref.allocate(int(s[8:]))
        elif s.startswith("cd"):
            ref.compile_default(int(s[2:]))
        else:
            raise ValueError(s)


def has_swaps(ref):
    return any(k == 0 for k, _ in ref.instructions())


def write(name, meta, arrays):
    arrays = dict(arrays)
    arrays["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrays)


QCX_CASES = [("mixed8", "raw"), ("mixed8", "sched_fuse5"), ("qft9", "raw"), ("qft9", "sched_group5_alloc2"),
             ("mixed11_dense4", "raw"), ("mixed11_dense4", "sched_fuse3_group6"), ("qulacs7", "alloc0"),
             ("brick12", "sched_group5_alloc2")]


def circuits():
    return {
        "mixed8": (8, W.random_mixed(8, 80, 3)),
        "layered10": (10, W.layered(10, 4, 5)),
        "qft9": (9, W.qft(9)),
        "brick12": (12, W.brickwork(12, 8, 2, longrange_every=4, longrange_offset=5)),
        "qulacs7": (7, W.qulacs_bench(7)),
        "mixed11_dense4": (11, W.random_mixed(11, 60, 7, max_dense=4)),
    }


def write_qcx():
    """Kate's Circuit::exportqcx (Circuit.hpp:763-774) of circuits after
    the named passes: kate_qcx_<circuit>_<pipeline>.qcx, byte for byte."""
    cs = circuits()
    for cname, pname in QCX_CASES:
        n, gates = cs[cname]
        ref = O.KateRef(n, gates)
        apply(ref, PIPELINES[pname])
        ref.export_qcx(os.path.join(OUT, f"kate_qcx_{cname}_{pname}.qcx"))
    return len(QCX_CASES)


def main():
    if "--qcx-only" in sys.argv:
        print(f"wrote {write_qcx()} QCX fixtures to {OUT}")
        return
    circuits_ = {
        "mixed8": (8, W.random_mixed(8, 80, 3)),
        "layered10": (10, W.layered(10, 4, 5)),
        "qft9": (9, W.qft(9)),
        "brick12": (12, W.brickwork(12, 8, 2, longrange_every=4, longrange_offset=5)),
        "qulacs7": (7, W.qulacs_bench(7)),
        "mixed11_dense4": (11, W.random_mixed(11, 60, 7, max_dense=4)),
    }
    assert circuits_.keys() == circuits().keys()
    written = write_qcx()
    for cname, (n, gates) in circuits_.items():
        for pname, steps in PIPELINES.items():
            ref = O.KateRef(n, gates)
            apply(ref, steps)
            out = ref.gates()
            arrays = {}
            arrays.update(pack_gates("in_", gates))
            arrays.update(pack_gates("out_", [(g[0], g[1], g[2], g[3]) for g in out], tags=[g[4] for g in out]))
            compiled = bool(ref.instructions())
            if compiled:
                arrays.update(pack_program("prog_", ref.groups(), ref.instructions(), ref.perms()))
            elif ref.groups():
                arrays.update(pack_program("prog_", ref.groups(), [], ([], [])))
            meta = {"circuit": cname, "n": n, "pipeline": steps, "compiled": compiled,
                    "grouped": bool(ref.groups()), "simulated": False}
            # Kate's CPU simulator ignores swap instructions, so only swap-free
            # programs have a meaningful Kate state.
            if not has_swaps(ref):
                st, pr = ref.simulate()
                arrays["state"] = st
                arrays["proba"] = pr
                meta["simulated"] = True
            write(f"kate_{cname}_{pname}", meta, arrays)
            written += 1

    # product-state inputs (execute(proba_state&), CPUcode/simulator.hpp:139-162)
    rng = np.random.default_rng(11)
    for cname, n, gates, tp in [
        ("qft9_basis", 9, W.qft(9), W.basis_input(9, 0b101101011)),
        ("mixed8_product", 8, W.random_mixed(8, 80, 3), np.column_stack([rng.uniform(0, 1, 8), rng.uniform(0, 6, 8)])),
    ]:
        for pname, steps in (("raw", []), ("alloc0", ["allocate0"])):
            ref = O.KateRef(n, gates)
            apply(ref, steps)
            st, pr = ref.simulate(tp)
            arrays = pack_gates("in_", gates)
            arrays.update(pack_gates("out_", [(g[0], g[1], g[2], g[3]) for g in ref.gates()]))
            if ref.instructions():
                arrays.update(pack_program("prog_", ref.groups(), ref.instructions(), ref.perms()))
            arrays.update({"state": st, "proba": pr, "proba_in": tp})
            write(f"kate_{cname}_{pname}", {"circuit": cname, "n": n, "pipeline": steps,
                                            "compiled": bool(ref.instructions()), "grouped": bool(ref.groups()),
                                            "simulated": True, "product_input": True}, arrays)
            written += 1

    # config 1: 20 qubits, 10 layers of RX/RZ + CNOT ladder, raw program
    n = 20
    gates = W.layered(n, 10, 1)
    ref = O.KateRef(n, gates)
    st, pr = ref.simulate()
    idx = np.sort(np.random.default_rng(2024).choice(1 << n, size=4096, replace=False))
    arrays = pack_gates("in_", gates)
    arrays.update({"sample_index": idx.astype(np.int64), "sample_amp": st[idx], "proba": pr,
                   "sum": np.array(st.sum()), "norm": np.array(np.vdot(st, st).real),
                   "zi": np.array([np.sum(np.abs(st) ** 2 * (1 - 2 * ((np.arange(1 << n) >> q) & 1))) for q in range(n)])})
    write("kate_config1_layered20", {"circuit": "layered20", "n": n, "pipeline": [], "compiled": False,
                                     "simulated": True, "sampled": True}, arrays)
    written += 1

    # CPUmergeGate on runs of a mixed circuit
    gates = W.random_mixed(6, 30, 5, max_dense=3)
    ref = O.KateRef(6, gates)
    for a, b in [(0, 4), (3, 9), (10, 14), (20, 30)]:
        qs, m = ref.merge(a, b)
        arrays = pack_gates("in_", gates[a:b])
        arrays.update({"qubits": np.array(qs, dtype=np.int32), "mat": m})
        write(f"kate_merge_{a}_{b}", {"n": 6, "first": a, "last": b}, arrays)
        written += 1
    print(f"wrote {written} fixtures to {OUT}")


if __name__ == "__main__":
    main()
