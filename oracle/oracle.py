"""ctypes wrappers for the oracle -- TEST INFRASTRUCTURE ONLY.

* ``liboracle.so`` (kate_oracle.c): plain-C restatement of Kate's CPU
  simulate path; see kate_oracle.h for the reference lines it follows.
* ``_ref/libkateref.so`` (kate_ref_harness.cpp): Kate's own CPU build, present
  only where the reference checkout was available at build time.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  Gates are ``(id, qubits, optarg, mat)`` records as produced by
``kate_amd.workloads``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
U64P = C.POINTER(C.c_uint64)
DP = C.POINTER(C.c_double)


def _p(a, t):
    return a.ctypes.data_as(t)


class _Gates(C.Structure):
    _fields_ = [("ngates", C.c_int32), ("id", I32P), ("nq", I32P), ("qoff", I32P), ("qubits", I32P),
                ("optarg", I32P), ("moff", I64P), ("mat", DP)]


class _Program(C.Structure):
    _fields_ = [("nqubits", C.c_int32), ("gates", _Gates), ("ngroups", C.c_int32), ("group_end", I32P),
                ("group_mask", U64P), ("ninstr", C.c_int32), ("instr_kind", I32P), ("instr_npairs", I32P),
                ("instr_poff", I32P), ("pairs", I32P)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _lib = C.CDLL(path)
        _lib.orc_run.argtypes = [C.POINTER(_Program), DP, C.c_int]
        _lib.orc_merge.argtypes = [C.POINTER(_Gates), C.c_int, C.c_int, I32P, DP]
        _lib.orc_init_basis.argtypes = [C.c_int, DP]
        _lib.orc_init_product.argtypes = [C.c_int, DP, I32P, DP]
        _lib.orc_measure.argtypes = [C.c_int, DP, I32P, DP]
        _lib.orc_expect_z.argtypes = [C.c_int, DP, C.c_int, U64P, DP]
        _lib.orc_swap_bits.argtypes = [C.c_int, DP, C.c_int, C.c_int]
        _lib.orc_apply_gate.argtypes = [C.POINTER(_Gates), C.c_int, C.c_int, DP, I32P]
    return _lib


class FlatGates:
    """Gate list flattened into the arrays orc_gates points at."""

    def __init__(self, gates: Sequence):
        n = len(gates)
        self.id = np.array([g.id for g in gates], dtype=np.int32)
        self.nq = np.array([len(g.qubits) for g in gates], dtype=np.int32)
        self.qoff = np.zeros(n, dtype=np.int32)
        self.qoff[1:] = np.cumsum(self.nq)[:-1] if n else []
        self.qubits = np.array([q for g in gates for q in g.qubits] or [0], dtype=np.int32)
        self.optarg = np.array([g.optarg for g in gates], dtype=np.int32)
        self.moff = np.full(n, -1, dtype=np.int64)
        mats = []
        off = 0
        for i, g in enumerate(gates):
            if g.id == 0:
                self.moff[i] = off
                m = np.ascontiguousarray(g.mat, dtype=np.complex128).view(np.float64).ravel()
                mats.append(m)
                off += m.size
        self.mat = np.concatenate(mats) if mats else np.zeros(1)
        self.n = n

    def struct(self) -> _Gates:
        return _Gates(self.n, _p(self.id, I32P), _p(self.nq, I32P), _p(self.qoff, I32P), _p(self.qubits, I32P),
                      _p(self.optarg, I32P), _p(self.moff, I64P), _p(self.mat, DP))


@dataclass
class Compiled:
    """Kate's program format (backendagnosticcode/QuantumCircuit.hpp:12-17)."""
    nqubits: int
    gates: list
    groups: List[Tuple[int, int]] = field(default_factory=list)  # (end, qubit mask)
    instructions: List[Tuple[int, List[int]]] = field(default_factory=list)
    initial_permutation: List[int] = field(default_factory=list)
    final_inverse_permutation: List[int] = field(default_factory=list)

    @staticmethod
    def naive(nqubits: int, gates: Sequence) -> "Compiled":
        """compileDefault(0) (QuantumCircuit.hpp:126-204): one group per gate,
        identity permutations, no swaps."""
        groups = []
        for i, g in enumerate(gates):
            m = 0
            for q in g.qubits:
                m |= 1 << q
            groups.append((i + 1, m))
        return Compiled(nqubits, list(gates), groups, [(1, [])] * len(gates), list(range(nqubits)),
                        list(range(nqubits)))


def run(prog: Compiled, proba_in: Optional[np.ndarray] = None, nthreads: int = 0) -> np.ndarray:
    """CPUSimulator::execute restated; returns amplitudes in virtual order."""
    L = lib()
    n = prog.nqubits
    fg = FlatGates(prog.gates)
    gend = np.array([e for e, _ in prog.groups] or [0], dtype=np.int32)
    gmask = np.array([m for _, m in prog.groups] or [0], dtype=np.uint64)
    kinds = np.array([k for k, _ in prog.instructions] or [0], dtype=np.int32)
    npairs = np.array([len(p) // 2 for _, p in prog.instructions] or [0], dtype=np.int32)
    poff = np.zeros(max(1, len(prog.instructions)), dtype=np.int32)
    if len(prog.instructions) > 1:
        poff[1:] = np.cumsum(npairs * 2)[:-1]
    pairs = np.array([x for _, p in prog.instructions for x in p] or [0], dtype=np.int32)
    p = _Program(n, fg.struct(), len(prog.groups), _p(gend, I32P), _p(gmask, U64P), len(prog.instructions),
                 _p(kinds, I32P), _p(npairs, I32P), _p(poff, I32P), _p(pairs, I32P))
    st = np.zeros(1 << n, dtype=np.complex128)
    sv = st.view(np.float64)
    if proba_in is None:
        L.orc_init_basis(n, _p(sv, DP))
    else:
        tp = np.ascontiguousarray(proba_in, dtype=np.float64).ravel()
        perm = np.array(prog.initial_permutation, dtype=np.int32)
        L.orc_init_product(n, _p(tp, DP), _p(perm, I32P), _p(sv, DP))
    nthreads = nthreads or min(16, len(os.sched_getaffinity(0)))
    rc = L.orc_run(C.byref(p), _p(sv, DP), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle run failed ({rc})")
    return st


def run_groups(prog: Compiled, state: np.ndarray, first: int, last: int, nthreads: int = 0):
    """Runs compute groups [first, last) of a swap-free program on `state`
    in place (the slice of CPUSimulator::execute's instruction loop that
    computes those groups)."""
    if any(k != 1 for k, _ in prog.instructions):
        raise ValueError("run_groups takes swap-free programs")
    g0 = prog.groups[first - 1][0] if first > 0 else 0
    g1 = prog.groups[last - 1][0]
    sub = Compiled(prog.nqubits, list(prog.gates[g0:g1]),
                   [(e - g0, m) for e, m in prog.groups[first:last]], [(1, [])] * (last - first),
                   prog.initial_permutation, prog.final_inverse_permutation)
    fg = FlatGates(sub.gates)
    gend = np.array([e for e, _ in sub.groups], dtype=np.int32)
    gmask = np.array([m for _, m in sub.groups], dtype=np.uint64)
    kinds = np.ones(len(sub.groups), dtype=np.int32)
    zeros = np.zeros(len(sub.groups), dtype=np.int32)
    pairs = np.zeros(1, dtype=np.int32)
    p = _Program(sub.nqubits, fg.struct(), len(sub.groups), _p(gend, I32P), _p(gmask, U64P), len(sub.groups),
                 _p(kinds, I32P), _p(zeros, I32P), _p(zeros, I32P), _p(pairs, I32P))
    nthreads = nthreads or min(16, len(os.sched_getaffinity(0)))
    rc = lib().orc_run(C.byref(p), _p(state.view(np.float64), DP), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle run failed ({rc})")


def init_basis(nqubits: int) -> np.ndarray:
    """|0...0> (CPUcode/simulator.hpp initialize)."""
    st = np.zeros(1 << nqubits, dtype=np.complex128)
    lib().orc_init_basis(nqubits, _p(st.view(np.float64), DP))
    return st


def simulate(nqubits: int, gates: Sequence, proba_in: Optional[np.ndarray] = None, nthreads: int = 0) -> np.ndarray:
    return run(Compiled.naive(nqubits, gates), proba_in, nthreads)


def measure(nq: int, state: np.ndarray, final_inverse_permutation: Optional[Sequence[int]] = None) -> np.ndarray:
    """Kate's proba_state: (theta, phi) per qubit (CPUcode/simulator.hpp:163-197)."""
    finv = np.array(final_inverse_permutation if final_inverse_permutation is not None else range(nq), dtype=np.int32)
    out = np.zeros(2 * nq)
    st = np.ascontiguousarray(state, dtype=np.complex128)
    lib().orc_measure(nq, _p(st.view(np.float64), DP), _p(finv, I32P), _p(out, DP))
    return out.reshape(nq, 2)


def expect_z(nq: int, state: np.ndarray, masks: Sequence[int]) -> np.ndarray:
    m = np.array(masks, dtype=np.uint64)
    out = np.zeros(len(m))
    st = np.ascontiguousarray(state, dtype=np.complex128)
    lib().orc_expect_z(nq, _p(st.view(np.float64), DP), len(m), _p(m, U64P), _p(out, DP))
    return out


def merge(gates: Sequence) -> Tuple[Tuple[int, ...], np.ndarray]:
    """CPUmergeGate restated: (ascending qubits, fused matrix)."""
    fg = FlatGates(gates)
    s = fg.struct()
    qs = np.zeros(64, dtype=np.int32)
    cover = set(q for g in gates for q in g.qubits)
    k = len(cover)
    mat = np.zeros((1 << k) * (1 << k), dtype=np.complex128)
    kk = lib().orc_merge(C.byref(s), 0, len(gates), _p(qs, I32P), _p(mat.view(np.float64), DP))
    return tuple(int(x) for x in qs[:kk]), mat.reshape(1 << kk, 1 << kk)


def swap_bits(nq: int, state: np.ndarray, p1: int, p2: int) -> np.ndarray:
    st = np.array(state, dtype=np.complex128, copy=True)
    lib().orc_swap_bits(nq, _p(st.view(np.float64), DP), p1, p2)
    return st


# ---------------------------------------------------------------- reference ---

def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libkateref.so"))


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libkateref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libkateref.so is not built (make -C oracle ref)")
        r = C.CDLL(path)
        r.kref_new.restype = C.c_void_p
        r.kref_new.argtypes = [C.c_int]
        r.kref_free.argtypes = [C.c_void_p]
        r.kref_append.argtypes = [C.c_void_p, C.c_int, C.c_int, I32P, C.c_int, DP, C.c_double]
        for f in ("kref_schedule",):
            getattr(r, f).argtypes = [C.c_void_p]
        r.kref_fuse.argtypes = [C.c_void_p, C.c_int, C.c_double]
        r.kref_group.argtypes = [C.c_void_p, C.c_int]
        r.kref_compile_default.argtypes = [C.c_void_p, C.c_int]
        r.kref_allocate.argtypes = [C.c_void_p, C.c_int]
        r.kref_ngates.argtypes = [C.c_void_p]
        r.kref_gate.argtypes = [C.c_void_p, C.c_int, I32P, I32P, I32P, I32P, DP, DP]
        r.kref_ngroups.argtypes = [C.c_void_p]
        r.kref_groups.argtypes = [C.c_void_p, I32P, U64P]
        r.kref_ninstr.argtypes = [C.c_void_p]
        r.kref_instr.argtypes = [C.c_void_p, C.c_int, I32P, I32P]
        r.kref_perms.argtypes = [C.c_void_p, I32P, I32P]
        r.kref_simulate.argtypes = [C.c_void_p, DP, DP, DP]
        r.kref_merge.argtypes = [C.c_void_p, C.c_int, C.c_int, I32P, DP]
        r.kref_time_execute.argtypes = [C.c_void_p, DP]
        r.kref_export_qcx.argtypes = [C.c_void_p, C.c_char_p]
        r.kref_qft.restype = C.c_void_p
        r.kref_qft.argtypes = [C.c_int]
        r.kref_qulacs.restype = C.c_void_p
        r.kref_qulacs.argtypes = [C.c_int]
        _ref = r
    return _ref


class KateRef:
    """Kate's own Circuit + CPUSimulator (CPU build of the reference)."""

    def __init__(self, nqubits: int, gates: Sequence = (), handle=None):
        self.r = ref()
        self.n = nqubits
        self.h = handle if handle is not None else self.r.kref_new(nqubits)
        for i, g in enumerate(gates):
            self.append(g, tag=float(i + 1))

    @classmethod
    def qft(cls, n):
        r = ref()
        return cls(n, handle=r.kref_qft(n))

    @classmethod
    def qulacs(cls, n):
        r = ref()
        return cls(n, handle=r.kref_qulacs(n))

    def __del__(self):
        try:
            self.r.kref_free(self.h)
        except Exception:
            pass

    def append(self, g, tag: float = 0.0):
        qs = np.array(g.qubits, dtype=np.int32)
        mat = (np.ascontiguousarray(g.mat, dtype=np.complex128).view(np.float64).ravel()
               if g.id == 0 else np.zeros(2))
        self.r.kref_append(self.h, g.id, len(qs), _p(qs, I32P), g.optarg, _p(mat, DP), tag)

    def schedule(self):
        self.r.kref_schedule(self.h)

    def fuse(self, k=5, mtm=1e-6):
        self.r.kref_fuse(self.h, k, mtm)

    def group(self, g=8):
        self.r.kref_group(self.h, g)

    def compile_default(self, gpulog2=0):
        self.r.kref_compile_default(self.h, gpulog2)

    def allocate(self, gpulog2=0):
        self.r.kref_allocate(self.h, gpulog2)

    def gates(self):
        """[(id, qubits, optarg, mat or None, tag)]"""
        out = []
        ids = np.zeros(1, dtype=np.int32)
        nq = np.zeros(1, dtype=np.int32)
        qs = np.zeros(64, dtype=np.int32)
        oa = np.zeros(1, dtype=np.int32)
        tag = np.zeros(1)
        for i in range(self.r.kref_ngates(self.h)):
            self.r.kref_gate(self.h, i, _p(ids, I32P), _p(nq, I32P), _p(qs, I32P), _p(oa, I32P), None,
                             _p(tag, DP))
            k = int(nq[0])
            m = None
            if ids[0] == 0:
                d = 1 << k
                mat = np.zeros(2 * d * d, dtype=np.float64)
                self.r.kref_gate(self.h, i, _p(ids, I32P), _p(nq, I32P), _p(qs, I32P), _p(oa, I32P), _p(mat, DP),
                                 _p(tag, DP))
                m = mat.view(np.complex128).reshape(d, d)
            out.append((int(ids[0]), tuple(int(x) for x in qs[:k]), int(oa[0]), m, float(tag[0])))
        return out

    def groups(self):
        n = self.r.kref_ngroups(self.h)
        e = np.zeros(max(1, n), dtype=np.int32)
        m = np.zeros(max(1, n), dtype=np.uint64)
        self.r.kref_groups(self.h, _p(e, I32P), _p(m, U64P))
        return [(int(e[i]), int(m[i])) for i in range(n)]

    def instructions(self):
        out = []
        kind = np.zeros(1, dtype=np.int32)
        pairs = np.zeros(4 * 64, dtype=np.int32)
        for i in range(self.r.kref_ninstr(self.h)):
            c = self.r.kref_instr(self.h, i, _p(kind, I32P), _p(pairs, I32P))
            out.append((int(kind[0]), [int(x) for x in pairs[:c]]))
        return out

    def perms(self):
        a = np.zeros(self.n, dtype=np.int32)
        b = np.zeros(self.n, dtype=np.int32)
        self.r.kref_perms(self.h, _p(a, I32P), _p(b, I32P))
        return [int(x) for x in a], [int(x) for x in b]

    def simulate(self, proba_in: Optional[np.ndarray] = None):
        """Returns (state in virtual order, proba_state (n,2))."""
        st = np.zeros(1 << self.n, dtype=np.complex128)
        pr = np.zeros(2 * self.n)
        pin = None
        if proba_in is not None:
            pin = np.ascontiguousarray(proba_in, dtype=np.float64).ravel()
        self.r.kref_simulate(self.h, _p(pin, DP) if pin is not None else None, _p(st.view(np.float64), DP),
                             _p(pr, DP))
        return st, pr.reshape(self.n, 2)

    def time_execute(self) -> float:
        """Wall ms of Kate's CPUSimulator::execute(true) (which prints its own
        init / computation / measurement split on stdout)."""
        ms = np.zeros(1)
        if self.r.kref_time_execute(self.h, _p(ms, DP)) != 0:
            raise RuntimeError("Kate's CPU simulator returned no proba_state")
        return float(ms[0])

    def export_qcx(self, path: str):
        """Kate's Circuit::exportqcx."""
        self.r.kref_export_qcx(self.h, path.encode())

    def merge(self, first, last):
        qs = np.zeros(64, dtype=np.int32)
        cover = set()
        for g in self.gates()[first:last]:
            cover.update(g[1])
        d = 1 << len(cover)
        mat = np.zeros(d * d, dtype=np.complex128)
        k = self.r.kref_merge(self.h, first, last, _p(qs, I32P), _p(mat.view(np.float64), DP))
        return tuple(int(x) for x in qs[:k]), mat.reshape(d, d)

    def compiled(self, gates_as) -> Compiled:
        """Kate's compiled program as an oracle Compiled (gates rebuilt with
        ``gates_as(id, qubits, optarg, mat)``)."""
        gl = [gates_as(i, q, o, m) for (i, q, o, m, _t) in self.gates()]
        ip, fi = self.perms()
        return Compiled(self.n, gl, self.groups(), self.instructions(), ip, fi)
