"""Packing of gate lists and compiled programs into .npz fixtures (data only)."""
from __future__ import annotations

import numpy as np

from kate_amd.workloads import Gate


def pack_gates(prefix: str, gates, tags=None) -> dict:
    ids = np.array([g[0] if isinstance(g, tuple) else g.id for g in gates], dtype=np.int32)
    qs = [tuple(g[1] if isinstance(g, tuple) else g.qubits) for g in gates]
    oa = np.array([g[2] if isinstance(g, tuple) else g.optarg for g in gates], dtype=np.int32)
    mats = [g[3] if isinstance(g, tuple) else g.mat for g in gates]
    nq = np.array([len(q) for q in qs], dtype=np.int32)
    flatq = np.array([x for q in qs for x in q], dtype=np.int32)
    moff = np.full(len(gates), -1, dtype=np.int64)
    pool = []
    off = 0
    for i, m in enumerate(mats):
        if ids[i] == 0:
            moff[i] = off
            pool.append(np.asarray(m, dtype=np.complex128).ravel())
            off += pool[-1].size
    out = {f"{prefix}ids": ids, f"{prefix}nq": nq, f"{prefix}qubits": flatq, f"{prefix}optarg": oa,
           f"{prefix}moff": moff,
           f"{prefix}mats": np.concatenate(pool) if pool else np.zeros(0, dtype=np.complex128)}
    if tags is not None:
        out[f"{prefix}tags"] = np.array(tags, dtype=np.float64)
    return out


def unpack_gates(z, prefix: str):
    ids, nq, flatq, oa, moff, mats = (z[f"{prefix}{k}"] for k in ("ids", "nq", "qubits", "optarg", "moff", "mats"))
    out = []
    p = 0
    for i in range(len(ids)):
        k = int(nq[i])
        q = tuple(int(x) for x in flatq[p:p + k])
        p += k
        m = None
        if ids[i] == 0:
            d = 1 << k
            m = mats[int(moff[i]):int(moff[i]) + d * d].reshape(d, d)
        out.append(Gate(int(ids[i]), q, int(oa[i]), m))
    return out


def pack_program(prefix: str, groups, instructions, perms) -> dict:
    ends = np.array([e for e, _ in groups], dtype=np.int32)
    masks = np.array([m for _, m in groups], dtype=np.uint64)
    kinds = np.array([k for k, _ in instructions], dtype=np.int32)
    plen = np.array([len(p) for _, p in instructions], dtype=np.int32)
    pairs = np.array([x for _, p in instructions for x in p], dtype=np.int32)
    ip, fi = perms
    return {f"{prefix}group_end": ends, f"{prefix}group_mask": masks, f"{prefix}instr_kind": kinds,
            f"{prefix}instr_len": plen, f"{prefix}pairs": pairs,
            f"{prefix}initial": np.array(ip, dtype=np.int32), f"{prefix}final_inverse": np.array(fi, dtype=np.int32)}


def unpack_program(z, prefix: str):
    groups = [(int(e), int(m)) for e, m in zip(z[f"{prefix}group_end"], z[f"{prefix}group_mask"])]
    instr = []
    p = 0
    for k, ln in zip(z[f"{prefix}instr_kind"], z[f"{prefix}instr_len"]):
        instr.append((int(k), [int(x) for x in z[f"{prefix}pairs"][p:p + int(ln)]]))
        p += int(ln)
    perms = ([int(x) for x in z[f"{prefix}initial"]], [int(x) for x in z[f"{prefix}final_inverse"]])
    return groups, instr, perms
