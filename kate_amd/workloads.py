"""Synthetic circuit generators for the simulate path (numpy only).

A gate is ``Gate(id, qubits, optarg, mat)`` with Kate's numbering and qubit
order (backendagnosticcode/agnosticpreprocessor.hpp:28-33):

* ``DENSE`` (0): ``mat`` is 2^k x 2^k complex128, row-major; matrix-index bit i
  is qubit ``qubits[i]`` (basic_host_types.hpp:229-245, GateComputing.hpp:103-125).
* ``H`` (2) on ``(q,)``.
* ``CNOT`` (3) on ``(target, control)`` -- Kate's ``qbits[1]`` is the control
  (CPUcode/GateComputing.hpp:46-64), so "CNOT i -> i+1" is ``(i+1, i)``.
* ``CRK`` (4) on ``(a, b)``, phase e^{2 pi i / 2^optarg} on |11>.
* ``TOFFOLI`` (5) on ``(c0, c1, target)``.

Kate computes no RX/RZ (ids 6/7 are defined but never evaluated), so the
rotation layers are emitted as dense 2x2 gates, exactly as a Kate user would
have to write them.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

DENSE, H, CNOT, CRK, TOFFOLI = 0, 2, 3, 4, 5


@dataclass(frozen=True)
class Gate:
    id: int
    qubits: Tuple[int, ...]
    optarg: int = 0
    mat: Optional[np.ndarray] = None

    def __post_init__(self):
        if len(set(self.qubits)) != len(self.qubits):
            raise ValueError(f"a qubit is present twice in {self.qubits}")
        if self.id == DENSE:
            if self.mat is None or self.mat.shape != (1 << len(self.qubits),) * 2:
                raise ValueError("dense gate needs a 2^k x 2^k matrix")


def rx(theta: float) -> np.ndarray:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def rz(phi: float) -> np.ndarray:
    return np.array([[complex(math.cos(-phi / 2), math.sin(-phi / 2)), 0],
                     [0, complex(math.cos(phi / 2), math.sin(phi / 2))]], dtype=np.complex128)


def layered(n: int, layers: int, seed: int) -> List[Gate]:
    """RX(theta) then RZ(phi) on every qubit, then the CNOT ladder i -> i+1.

    theta, phi ~ U[0, 2 pi) from ``seed``.  10*(2n + n-1) gates for 10 layers.
    """
    rng = np.random.default_rng(seed)
    gates: List[Gate] = []
    for _ in range(layers):
        ang = rng.uniform(0.0, 2 * math.pi, size=(n, 2))
        for q in range(n):
            gates.append(Gate(DENSE, (q,), 0, rx(float(ang[q, 0]))))
            gates.append(Gate(DENSE, (q,), 0, rz(float(ang[q, 1]))))
        for i in range(n - 1):
            gates.append(Gate(CNOT, (i + 1, i)))
    return gates


def haar_unitary(rng: np.random.Generator, dim: int) -> np.ndarray:
    """Haar-random U(dim): QR of a complex Ginibre matrix with the phases of
    R's diagonal moved into Q (Mezzadri, Notices AMS 54 (2007) 592)."""
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return (q * (d / np.abs(d))).astype(np.complex128)


def haar_layers(n: int, layers: int, seed: int) -> List[Gate]:
    """A Haar-random U(2) on every qubit, then the CNOT ladder i -> i+1, per
    layer: general one-qubit gates (no RX/RZ structure)."""
    rng = np.random.default_rng(seed)
    gates: List[Gate] = []
    for _ in range(layers):
        for q in range(n):
            gates.append(Gate(DENSE, (q,), 0, haar_unitary(rng, 2)))
        for i in range(n - 1):
            gates.append(Gate(CNOT, (i + 1, i)))
    return gates


def brickwork(n: int, depth: int, seed: int, longrange_every: int = 4, longrange_offset: int = 16) -> List[Gate]:
    """Haar 2-qubit brickwork: even layers pair (2j, 2j+1), odd layers
    (2j+1, 2j+2).  Every ``longrange_every``-th layer (1-based) also applies
    Haar gates on disjoint random pairs (q, q+offset mod n)."""
    rng = np.random.default_rng(seed)
    gates: List[Gate] = []
    for layer in range(depth):
        start = layer % 2
        for a in range(start, n - 1, 2):
            gates.append(Gate(DENSE, (a, a + 1), 0, haar_unitary(rng, 4)))
        if longrange_every and (layer + 1) % longrange_every == 0 and n > longrange_offset:
            used = set()
            for q in rng.permutation(n):
                p = (int(q) + longrange_offset) % n
                q = int(q)
                if q in used or p in used or p == q:
                    continue
                used.update((q, p))
                gates.append(Gate(DENSE, (q, p), 0, haar_unitary(rng, 4)))
    return gates


def qft(n: int) -> List[Gate]:
    """Kate's QFT generator (BenchCircuits.hpp:44-58): H then CRk(k=j-i+1) on
    (j, i), no final swaps (the output is bit-reversed)."""
    gates: List[Gate] = []
    for i in range(n):
        gates.append(Gate(H, (i,)))
        for j in range(i + 1, n):
            gates.append(Gate(CRK, (j, i), j - i + 1))
    return gates


def qulacs_bench(n: int) -> List[Gate]:
    """Kate's QulacsBench generator (BenchCircuits.hpp:8-42): 41 n gates."""
    gates: List[Gate] = []
    for i in range(n):
        gates += [Gate(H, (i,)), Gate(H, (i,))]
    for i in range(n):
        gates.append(Gate(CNOT, (i, (i + 1) % n)))
    for _ in range(9):
        for i in range(n):
            gates += [Gate(H, (i,)), Gate(H, (i,)), Gate(H, (i,))]
        for i in range(n):
            gates.append(Gate(CNOT, (i, (i + 1) % n)))
    for i in range(n):
        gates += [Gate(H, (i,)), Gate(H, (i,))]
    return gates


def random_mixed(n: int, ngates: int, seed: int, max_dense: int = 3) -> List[Gate]:
    """Every gate kind Kate computes, on random qubits, with Haar dense gates."""
    rng = np.random.default_rng(seed)
    gates: List[Gate] = []
    for _ in range(ngates):
        kind = int(rng.integers(0, 5))
        if kind == 0:
            k = int(rng.integers(1, min(max_dense, n) + 1))
            qs = tuple(int(x) for x in rng.choice(n, size=k, replace=False))
            gates.append(Gate(DENSE, qs, 0, haar_unitary(rng, 1 << k)))
        elif kind == 1:
            gates.append(Gate(H, (int(rng.integers(0, n)),)))
        elif kind == 2 and n >= 2:
            qs = tuple(int(x) for x in rng.choice(n, size=2, replace=False))
            gates.append(Gate(CNOT, qs))
        elif kind == 3 and n >= 2:
            qs = tuple(int(x) for x in rng.choice(n, size=2, replace=False))
            gates.append(Gate(CRK, qs, int(rng.integers(1, 8))))
        elif n >= 3:
            qs = tuple(int(x) for x in rng.choice(n, size=3, replace=False))
            gates.append(Gate(TOFFOLI, qs))
        else:
            gates.append(Gate(H, (int(rng.integers(0, n)),)))
    return gates


def basis_input(n: int, x: int) -> np.ndarray:
    """(theta, phi) pairs of Kate's proba_state for the basis state |x>:
    theta = 1 puts a qubit in |1> (CPUcode/simulator.hpp:139-162)."""
    tp = np.zeros((n, 2))
    for q in range(n):
        tp[q, 0] = float((x >> q) & 1)
    return tp


def zstring_masks(n: int, count: int, seed: int) -> np.ndarray:
    """``count`` seeded Pauli-Z strings (non-empty qubit bitmasks)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        m = int(rng.integers(1, 1 << min(n, 62))) if n < 63 else int(rng.integers(1, 1 << 62))
        out.append(m)
    return np.array(out, dtype=np.uint64)


def gate_counts(gates: Sequence[Gate]) -> dict:
    out: dict = {}
    for g in gates:
        out[g.id] = out.get(g.id, 0) + 1
    return out


def haar_unitaries(rng: np.random.Generator, count: int, dim: int) -> np.ndarray:
    """``count`` Haar-random U(dim) matrices at once (Mezzadri's QR recipe)."""
    z = (rng.standard_normal((count, dim, dim)) + 1j * rng.standard_normal((count, dim, dim))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    return (q * (d / np.abs(d))[:, None, :]).astype(np.complex128)


def batch_workload(count: int, seed: int, nmin: int = 10, nmax: int = 20, dmin: int = 8, dmax: int = 32):
    """Config-C4 style batch: qubit counts round(exp(U[ln nmin, ln nmax])),
    depths uniform in [dmin, dmax], Haar 2-qubit brickwork (no long-range
    layers).  Returns [(n, gates)]."""
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=count))).astype(int)
    depths = rng.integers(dmin, dmax + 1, size=count)
    need = int(sum(d * (n // 2) for n, d in zip(ns, depths)) + count)
    mats = haar_unitaries(rng, need, 4)
    out, k = [], 0
    for n, d in zip(ns, depths):
        gates = []
        for layer in range(int(d)):
            for a in range(layer % 2, int(n) - 1, 2):
                gates.append(Gate(DENSE, (a, a + 1), 0, mats[k]))
                k += 1
        out.append((int(n), gates))
    return out
