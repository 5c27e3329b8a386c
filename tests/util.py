"""Helpers shared by the tests."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

from fixture_io import unpack_gates, unpack_program
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixtures(pattern="kate_*.npz"):
    return sorted(glob.glob(os.path.join(GOLDEN, pattern)))


def load(path):
    z = np.load(path, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return z, meta


def compiled_from_circuit(c) -> O.Compiled:
    """The program a kate_amd.Circuit holds, as the oracle's Compiled
    (naively compiled first when it has no instructions, as the simulator does)."""
    gates = c.gates()
    ins = c.instructions()
    if not ins:
        groups = c.groups()
        if groups:
            ip = list(range(c.n))
            return O.Compiled(c.n, gates, groups, [(1, [])] * len(groups), ip, ip)
        return O.Compiled.naive(c.n, gates)
    ip, fi = c.permutations()
    return O.Compiled(c.n, gates, c.groups(), ins, ip, fi)


def compiled_from_fixture(z, meta) -> O.Compiled:
    gates = unpack_gates(z, "out_")
    n = meta["n"]
    if meta.get("compiled"):
        groups, ins, perms = unpack_program(z, "prog_")
        return O.Compiled(n, gates, groups, ins, perms[0], perms[1])
    if meta.get("grouped"):
        groups, _, _ = unpack_program(z, "prog_")
        ip = list(range(n))
        return O.Compiled(n, gates, groups, [(1, [])] * len(groups), ip, ip)
    return O.Compiled.naive(n, gates)


def to_real_order(state, final_inverse):
    """Virtual-order amplitudes -> real qubit order."""
    n = len(final_inverse)
    idx = np.arange(1 << n)
    real = np.zeros_like(idx)
    for p, r in enumerate(final_inverse):
        real |= ((idx >> p) & 1) << r
    out = np.empty_like(state)
    out[real] = state
    return out


def dense_unitary(n, gates):
    """Brute force: the 2^n x 2^n matrix of a gate list, built with numpy only
    (Kate's conventions: little-endian index, matrix bit i <-> qubits[i])."""
    dim = 1 << n
    U = np.eye(dim, dtype=np.complex128)
    for g in gates:
        G = np.zeros((dim, dim), dtype=np.complex128)
        k = len(g.qubits)
        if g.id == 0:
            M = g.mat
        elif g.id == 2:
            M = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2)
        elif g.id == 3:  # (target, control): index bit0 = target, bit1 = control
            M = np.eye(4, dtype=np.complex128)[[0, 1, 3, 2]]
        elif g.id == 4:
            M = np.diag([1, 1, 1, np.exp(2j * np.pi / (1 << g.optarg))])
        elif g.id == 5:  # (c0, c1, target): bit2 = target
            perm = list(range(8))
            perm[3], perm[7] = 7, 3
            M = np.eye(8, dtype=np.complex128)[perm]
        else:
            raise ValueError(g.id)
        for x in range(dim):
            sub = 0
            for i, q in enumerate(g.qubits):
                sub |= ((x >> q) & 1) << i
            base = x
            for q in g.qubits:
                base &= ~(1 << q)
            for r in range(1 << k):
                y = base
                for i, q in enumerate(g.qubits):
                    y |= ((r >> i) & 1) << q
                G[y, x] += M[r, sub]
        U = G @ U
    return U


def fidelity(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)
