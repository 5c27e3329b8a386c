"""kate_amd -- MI355X state-vector simulator with Kate's API (Python binding).

ctypes over the C ABI of ``libkate_hip.so`` (include/kate_hip.h).  The
classes mirror Kate's C++ API: ``Circuit`` (backendagnosticcode/Circuit.hpp:
gateScheduling / gateFusion / gateGrouping / allocate / compileDefault) and
``Simulator`` (HIPcode/simulator.hpp: execute, proba_state).  There is no CPU
path: without the built library or a GPU every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from .workloads import Gate  # noqa: F401  (re-exported)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkate_hip.so")

KATE_C128, KATE_C64 = 0, 1
MAX_GATE_QUBITS = 16

_lib = None


class KateError(RuntimeError):
    pass


class _GateRec(C.Structure):
    _fields_ = [("id", C.c_int32), ("nq", C.c_int32), ("qubits", C.c_int32 * MAX_GATE_QUBITS),
                ("optarg", C.c_int32), ("reserved", C.c_int32), ("matrix_offset", C.c_int64)]


class Options(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("tile_qubits", C.c_int32), ("low_qubits", C.c_int32),
                ("use_mfma", C.c_int32), ("batch_wgs_per_cu", C.c_int32), ("device", C.c_int32),
                ("micro_bits", C.c_int32), ("plan_flags", C.c_int32)]


PLAN_NO_EULER, PLAN_NO_DIRECT, PLAN_NO_SUPPORT, PLAN_NO_DEFER, PLAN_NO_SPLIT, PLAN_NO_REORDER = 1, 2, 4, 8, 16, 32
PLAN_NO_FACTOR = 64
PLAN_NO_SCATTER = 128
PLAN_NO_RELABEL = 256
PLAN_NO_MATRIX5 = 512


def _plan_flags(euler: bool, direct: bool, support: bool = True, defer: bool = True, split: bool = True,
                reorder: bool = True, factor: bool = True, scatter: bool = True, relabel: bool = True,
                matrix5: bool = True) -> int:
    return ((0 if euler else PLAN_NO_EULER) | (0 if direct else PLAN_NO_DIRECT) | (0 if support else PLAN_NO_SUPPORT)
            | (0 if defer else PLAN_NO_DEFER) | (0 if split else PLAN_NO_SPLIT)
            | (0 if reorder else PLAN_NO_REORDER) | (0 if factor else PLAN_NO_FACTOR)
            | (0 if scatter else PLAN_NO_SCATTER) | (0 if relabel else PLAN_NO_RELABEL)
            | (0 if matrix5 else PLAN_NO_MATRIX5))


class Timings(C.Structure):
    _fields_ = [("init_ms", C.c_double), ("compute_ms", C.c_double), ("measure_ms", C.c_double),
                ("passes", C.c_int32), ("remap_passes", C.c_int32), ("micro_steps", C.c_int32),
                ("gates", C.c_int32), ("bytes_per_pass", C.c_double)]


class ServerOptions(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("device", C.c_int32), ("streams", C.c_int32), ("max_active", C.c_int32),
                ("hbm_budget_bytes", C.c_int64), ("tile_qubits", C.c_int32), ("low_qubits", C.c_int32),
                ("use_mfma", C.c_int32), ("micro_bits", C.c_int32), ("reserved", C.c_int32 * 2)]


class RequestStats(C.Structure):
    _fields_ = [("submit_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
                ("latency_ms", C.c_double), ("passes", C.c_int32), ("nqubits", C.c_int32),
                ("footprint_bytes", C.c_int64)]


EXPORTS = [
    "kate_status_string", "kate_last_error", "kate_device_count", "kate_device_info",
    "kate_sim_create", "kate_sim_destroy", "kate_sim_init_basis", "kate_sim_init_product", "kate_sim_run",
    "kate_sim_measure_proba", "kate_sim_execute", "kate_sim_expect_z", "kate_sim_norm", "kate_sim_renormalize",
    "kate_sim_read_state", "kate_sim_write_state", "kate_sim_device_state", "kate_sim_last_timings",
    "kate_sim_synchronize", "kate_sim_stream", "kate_sim_describe", "kate_plan_describe", "kate_sim_run_timed",
    "kate_sim_overlap", "kate_sim_execute_timed", "kate_sim_copy_ceiling", "kate_state_expect_z_noalloc",
    "kate_circuit_create", "kate_circuit_destroy", "kate_circuit_append", "kate_circuit_schedule",
    "kate_circuit_fuse", "kate_circuit_group", "kate_circuit_allocate", "kate_circuit_compile_default",
    "kate_circuit_compile_opt", "kate_circuit_export_qcx", "kate_circuit_info", "kate_circuit_gate",
    "kate_circuit_groups", "kate_circuit_instruction", "kate_circuit_permutations", "kate_sim_create_from_circuit",
    "kate_circuit_plan", "kate_batch_create", "kate_batch_create_from_circuits", "kate_batch_destroy",
    "kate_batch_capture", "kate_batch_run", "kate_batch_info", "kate_batch_expect_z", "kate_batch_last_ms",
    "kate_server_create", "kate_server_destroy", "kate_server_submit", "kate_server_submit_circuit",
    "kate_server_wait", "kate_server_pending",
]

VP = C.c_void_p
I32P = C.POINTER(C.c_int32)
U64P = C.POINTER(C.c_uint64)
DP = C.POINTER(C.c_double)


def lib():
    """Load libkate_hip.so (raises KateError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise KateError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C kate_amd")
        L = C.CDLL(LIB_PATH)
        L.kate_status_string.restype = C.c_char_p
        L.kate_status_string.argtypes = [C.c_int]
        L.kate_last_error.restype = C.c_char_p
        L.kate_device_count.argtypes = [I32P]
        L.kate_device_info.argtypes = [C.c_char_p, C.c_size_t]
        L.kate_sim_create.argtypes = [VP, C.POINTER(Options), C.POINTER(VP)]
        L.kate_sim_destroy.argtypes = [VP]
        L.kate_sim_init_basis.argtypes = [VP, C.c_uint64]
        L.kate_sim_init_product.argtypes = [VP, DP]
        L.kate_sim_run.argtypes = [VP]
        L.kate_sim_measure_proba.argtypes = [VP, DP]
        L.kate_sim_execute.argtypes = [VP, DP, DP]
        L.kate_sim_expect_z.argtypes = [VP, U64P, C.c_int32, DP]
        L.kate_sim_norm.argtypes = [VP, DP]
        L.kate_sim_renormalize.argtypes = [VP]
        L.kate_sim_read_state.argtypes = [VP, VP, C.c_uint64, C.c_int32]
        L.kate_sim_write_state.argtypes = [VP, VP, C.c_uint64]
        L.kate_sim_device_state.argtypes = [VP, C.POINTER(VP)]
        L.kate_sim_last_timings.argtypes = [VP, C.POINTER(Timings)]
        L.kate_sim_synchronize.argtypes = [VP]
        L.kate_sim_stream.argtypes = [VP, C.POINTER(VP)]
        L.kate_sim_describe.argtypes = [VP, C.c_char_p, C.c_size_t]
        L.kate_circuit_create.argtypes = [C.c_int32, C.POINTER(VP)]
        L.kate_circuit_destroy.argtypes = [VP]
        L.kate_circuit_append.argtypes = [VP, C.c_int32, C.c_int32, I32P, C.c_int32, DP]
        L.kate_circuit_schedule.argtypes = [VP]
        L.kate_circuit_fuse.argtypes = [VP, C.c_int32, C.c_double]
        L.kate_circuit_group.argtypes = [VP, C.c_int32]
        L.kate_circuit_allocate.argtypes = [VP, C.c_int32, C.c_int32]
        L.kate_circuit_compile_default.argtypes = [VP, C.c_int32, C.c_int32]
        L.kate_circuit_compile_opt.argtypes = [VP, C.c_int32, C.c_double, C.c_int32, C.c_int32, C.c_int32]
        L.kate_circuit_export_qcx.argtypes = [VP, C.c_char_p]
        L.kate_circuit_info.argtypes = [VP, I32P, I32P, I32P, I32P]
        L.kate_circuit_gate.argtypes = [VP, C.c_int32, I32P, I32P, I32P, I32P, DP]
        L.kate_circuit_groups.argtypes = [VP, I32P, U64P]
        L.kate_circuit_instruction.argtypes = [VP, C.c_int32, I32P, I32P, C.c_int32, I32P]
        L.kate_circuit_permutations.argtypes = [VP, I32P, I32P]
        L.kate_sim_create_from_circuit.argtypes = [VP, C.POINTER(Options), C.POINTER(VP)]
        L.kate_sim_overlap.argtypes = [VP, VP, DP]
        L.kate_sim_run_timed.argtypes = [VP, C.POINTER(C.c_float), I32P, C.c_int32, I32P]
        L.kate_sim_execute_timed.argtypes = [VP, DP, DP, C.POINTER(C.c_float), U64P, C.c_int32, I32P]
        L.kate_sim_copy_ceiling.argtypes = [VP, C.POINTER(C.c_float), C.c_int32, I32P]
        L.kate_state_expect_z_noalloc.argtypes = [C.c_int32, C.c_int32, VP, C.c_int32, U64P, C.c_int32, VP, VP]
        L.kate_plan_describe.argtypes = [VP, C.POINTER(Options), I32P, I32P, C.c_char_p, C.c_size_t]
        L.kate_circuit_plan.argtypes = [VP, C.POINTER(Options), I32P, I32P, C.c_char_p, C.c_size_t]
        L.kate_batch_create.argtypes = [VP, C.c_int32, C.POINTER(Options), C.POINTER(VP)]
        L.kate_batch_create_from_circuits.argtypes = [C.POINTER(VP), C.c_int32, C.POINTER(Options), C.POINTER(VP)]
        L.kate_batch_destroy.argtypes = [VP]
        L.kate_batch_capture.argtypes = [VP]
        L.kate_batch_run.argtypes = [VP]
        L.kate_batch_info.argtypes = [VP, I32P, C.POINTER(C.c_int64), I32P, I32P, I32P]
        L.kate_batch_expect_z.argtypes = [VP, DP, C.c_int64, DP]
        L.kate_batch_last_ms.argtypes = [VP, C.POINTER(C.c_float)]
        L.kate_server_create.argtypes = [C.POINTER(ServerOptions), C.POINTER(VP)]
        L.kate_server_destroy.argtypes = [VP]
        L.kate_server_submit.argtypes = [VP, VP, DP, U64P, C.c_int32, C.POINTER(C.c_uint64)]
        L.kate_server_submit_circuit.argtypes = [VP, VP, DP, U64P, C.c_int32, C.POINTER(C.c_uint64)]
        L.kate_server_wait.argtypes = [VP, C.c_uint64, DP, C.c_int32, C.POINTER(RequestStats)]
        L.kate_server_pending.argtypes = [VP, I32P, I32P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().kate_last_error().decode(errors="replace")
        raise KateError(f"{what}: {lib().kate_status_string(rc).decode()} ({msg})")


def _p(a, t):
    return a.ctypes.data_as(t)


def device_count() -> int:
    n = C.c_int32(0)
    rc = lib().kate_device_count(C.byref(n))
    return int(n.value) if rc == 0 else 0


def device_info() -> str:
    buf = C.create_string_buffer(4096)
    _check(lib().kate_device_info(buf, 4096), "kate_device_info")
    return buf.value.decode()


class Circuit:
    """Kate::Circuit through the C ABI: build, run the passes, read back."""

    def __init__(self, nqubits: int, gates: Iterable = ()):
        self.n = nqubits
        h = VP()
        _check(lib().kate_circuit_create(nqubits, C.byref(h)), "kate_circuit_create")
        self._h = h
        for g in gates:
            self.append(g)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.kate_circuit_destroy(h)
            self._h = None

    def append(self, g) -> "Circuit":
        qs = np.array(g.qubits, dtype=np.int32)
        mat = None
        if g.id == 0:
            mat = np.ascontiguousarray(g.mat, dtype=np.complex128).view(np.float64).ravel()
        _check(lib().kate_circuit_append(self._h, g.id, len(qs), _p(qs, I32P), g.optarg,
                                         _p(mat, DP) if mat is not None else None), "kate_circuit_append")
        return self

    def extend(self, gates: Iterable) -> "Circuit":
        for g in gates:
            self.append(g)
        return self

    # passes (backendagnosticcode/Circuit.hpp)
    def schedule(self):
        _check(lib().kate_circuit_schedule(self._h), "gateScheduling")
        return self

    def fuse(self, max_qubits: int = 5, merge_time_matters: float = 1e-6):
        _check(lib().kate_circuit_fuse(self._h, max_qubits, merge_time_matters), "gateFusion")
        return self

    def group(self, group_qubits: int = 8):
        _check(lib().kate_circuit_group(self._h, group_qubits), "gateGrouping")
        return self

    def allocate(self, gpu_log2: int = 0, max_local: int = 300):
        _check(lib().kate_circuit_allocate(self._h, gpu_log2, max_local), "allocate")
        return self

    def compile_default(self, gpu_log2: int = 0, max_local: int = 300):
        _check(lib().kate_circuit_compile_default(self._h, gpu_log2, max_local), "compileDefault")
        return self

    def compile_opt(self, fuse_qubits=5, merge_time_matters=1e-6, group_qubits=10, gpu_log2=0, max_local=300):
        _check(lib().kate_circuit_compile_opt(self._h, fuse_qubits, merge_time_matters, group_qubits, gpu_log2,
                                              max_local), "compileOPT")
        return self

    def export_qcx(self, path: str):
        _check(lib().kate_circuit_export_qcx(self._h, path.encode()), "exportqcx")

    def info(self) -> Tuple[int, int, int, int]:
        a, b, c, d = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().kate_circuit_info(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)), "info")
        return a.value, b.value, c.value, d.value

    def gates(self) -> List[Gate]:
        out = []
        _, ng, _, _ = self.info()
        idv, nq, oa = C.c_int32(), C.c_int32(), C.c_int32()
        qs = np.zeros(MAX_GATE_QUBITS, dtype=np.int32)
        for i in range(ng):
            _check(lib().kate_circuit_gate(self._h, i, C.byref(idv), C.byref(nq), _p(qs, I32P), C.byref(oa), None),
                   "gate")
            k = nq.value
            mat = None
            if idv.value == 0:
                m = np.zeros(2 * (1 << (2 * k)), dtype=np.float64)
                _check(lib().kate_circuit_gate(self._h, i, C.byref(idv), C.byref(nq), _p(qs, I32P), C.byref(oa),
                                               _p(m, DP)), "gate")
                mat = m.view(np.complex128).reshape(1 << k, 1 << k)
            out.append(Gate(idv.value, tuple(int(x) for x in qs[:k]), oa.value, mat))
        return out

    def groups(self) -> List[Tuple[int, int]]:
        _, _, ngr, _ = self.info()
        e = np.zeros(max(1, ngr), dtype=np.int32)
        m = np.zeros(max(1, ngr), dtype=np.uint64)
        _check(lib().kate_circuit_groups(self._h, _p(e, I32P), _p(m, U64P)), "groups")
        return [(int(e[i]), int(m[i])) for i in range(ngr)]

    def instructions(self) -> List[Tuple[int, List[int]]]:
        _, _, _, ni = self.info()
        out = []
        kind, cnt = C.c_int32(), C.c_int32()
        buf = np.zeros(256, dtype=np.int32)
        for i in range(ni):
            _check(lib().kate_circuit_instruction(self._h, i, C.byref(kind), _p(buf, I32P), 256, C.byref(cnt)),
                   "instruction")
            out.append((kind.value, [int(x) for x in buf[:cnt.value]]))
        return out

    def plan(self, dtype: str = "c128", tile_qubits: int = 0, low_qubits: int = 0, mfma: int = 0,
             micro_bits: int = 0, euler: bool = True, direct: bool = True,
             split: bool = True, reorder: bool = True, factor: bool = True,
             scatter: bool = True, relabel: bool = True, matrix5: bool = True) -> Tuple[int, int, str]:
        """(gate passes, remap passes, text) of the device plan, without a GPU
        (a simulator plans the scatter pass only when the GPU has room for
        a second state buffer)."""
        opt = Options({"c128": KATE_C128, "c64": KATE_C64}[dtype], tile_qubits, low_qubits, mfma, 0, 0, micro_bits,
                      _plan_flags(euler, direct, split=split, reorder=reorder, factor=factor, scatter=scatter,
                                  relabel=relabel, matrix5=matrix5))
        p, r = C.c_int32(), C.c_int32()
        buf = C.create_string_buffer(1 << 20)
        _check(lib().kate_circuit_plan(self._h, C.byref(opt), C.byref(p), C.byref(r), buf, 1 << 20), "plan")
        return p.value, r.value, buf.value.decode()

    def permutations(self) -> Tuple[List[int], List[int]]:
        a = np.zeros(self.n, dtype=np.int32)
        b = np.zeros(self.n, dtype=np.int32)
        _check(lib().kate_circuit_permutations(self._h, _p(a, I32P), _p(b, I32P)), "permutations")
        ip, fi = [int(x) for x in a], [int(x) for x in b]
        _, _, _, ni = self.info()
        if ni == 0:
            return [], []
        return ip, fi


class Simulator:
    """Kate::Simulator on one GPU (HIPcode/simulator.hpp:220-543)."""

    def __init__(self, circuit: Circuit, dtype: str = "c128", tile_qubits: int = 0, low_qubits: int = 0,
                 device: int = 0, mfma: int = 0, micro_bits: int = 0, euler: bool = True, direct: bool = True,
                 support: bool = True, defer: bool = False, split: bool = True, reorder: bool = True,
                 factor: bool = True, scatter: bool = True, relabel: bool = True, matrix5: bool = True):
        """`defer=True` lets execute() leave the trailing passes that only
        permute amplitudes (CNOT layers at the end of the circuit) unrun:
        its proba_state reads through them and any later read of the state
        runs them first.  Off by default, as Kate's execute() leaves the
        finished state in memory.  `split=False` applies dense 2-qubit gates
        of the form CNOT x (A (x) B) as 4x4 products instead of as their
        factors; `reorder=False` plans the gates only in the order given;
        `factor=False` keeps Euler-form Kronecker steps unscaled;
        `scatter=False` runs the trailing permutation passes as planned
        instead of as one out-of-place pass; `relabel=False` keeps every
        pass in place on the virtual qubit order; `matrix5=False` keeps dense
        5-qubit gates of a complex64 state on the VALU instead of the matrix
        cores."""
        self.n = circuit.n
        self.dtype = {"c128": KATE_C128, "c64": KATE_C64}[dtype]
        self.amp = np.complex128 if self.dtype == KATE_C128 else np.complex64
        opt = Options(self.dtype, tile_qubits, low_qubits, mfma, 0, device, micro_bits,
                      _plan_flags(euler, direct, support, defer, split, reorder, factor, scatter, relabel, matrix5))
        h = VP()
        _check(lib().kate_sim_create_from_circuit(circuit._h, C.byref(opt), C.byref(h)), "kate_sim_create")
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.kate_sim_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def execute(self, proba_in: Optional[np.ndarray] = None) -> np.ndarray:
        """Kate's execute(): init (|0..0> or product state), run, proba_state (n, 2)."""
        out = np.zeros(2 * self.n)
        pin = None if proba_in is None else np.ascontiguousarray(proba_in, dtype=np.float64).ravel()
        _check(lib().kate_sim_execute(self._h, _p(pin, DP) if pin is not None else None, _p(out, DP)), "execute")
        return out.reshape(self.n, 2)

    def init_basis(self, index: int = 0):
        _check(lib().kate_sim_init_basis(self._h, index), "init_basis")

    def init_product(self, theta_phi: np.ndarray):
        tp = np.ascontiguousarray(theta_phi, dtype=np.float64).ravel()
        _check(lib().kate_sim_init_product(self._h, _p(tp, DP)), "init_product")

    def run(self):
        _check(lib().kate_sim_run(self._h), "run")

    def run_timed(self) -> Tuple[np.ndarray, np.ndarray]:
        """Run every pass with HIP events around each; (ms per pass, is-remap flags)."""
        cap = 1 << 16
        ms = np.zeros(cap, dtype=np.float32)
        rm = np.zeros(cap, dtype=np.int32)
        n = C.c_int32()
        _check(lib().kate_sim_run_timed(self._h, ms.ctypes.data_as(C.POINTER(C.c_float)), _p(rm, I32P), cap,
                                        C.byref(n)), "run_timed")
        return ms[:n.value].astype(np.float64), rm[:n.value].astype(bool)

    def execute_timed(self, proba_in: Optional[np.ndarray] = None) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """execute() with a HIP event after every pass: (proba_state, ms per
        pass, tiles launched per pass) for exactly the passes execute() runs."""
        out = np.zeros(2 * self.n)
        pin = None if proba_in is None else np.ascontiguousarray(proba_in, dtype=np.float64).ravel()
        cap = 1 << 16
        ms = np.zeros(cap, dtype=np.float32)
        tiles = np.zeros(cap, dtype=np.uint64)
        n = C.c_int32()
        _check(lib().kate_sim_execute_timed(self._h, _p(pin, DP) if pin is not None else None, _p(out, DP),
                                            ms.ctypes.data_as(C.POINTER(C.c_float)), _p(tiles, U64P), cap,
                                            C.byref(n)), "execute_timed")
        return out.reshape(self.n, 2), ms[:n.value].astype(np.float64), tiles[:n.value].copy()

    def copy_ceiling(self) -> np.ndarray:
        """Per pass: ms of a plain in-place copy of its tiles (same tile
        positions, no arithmetic): the HBM time its access pattern allows."""
        cap = 1 << 16
        ms = np.zeros(cap, dtype=np.float32)
        n = C.c_int32()
        _check(lib().kate_sim_copy_ceiling(self._h, ms.ctypes.data_as(C.POINTER(C.c_float)), cap, C.byref(n)),
               "copy_ceiling")
        return ms[:n.value].astype(np.float64)

    def compare(self, other: "Simulator") -> Tuple[complex, float, float, float]:
        """(<self|other>, <self|self>, <other|other>, max |self_x - other_x|),
        reduced in FP64 on the device."""
        out = np.zeros(5)
        _check(lib().kate_sim_overlap(self._h, other._h, _p(out, DP)), "overlap")
        return complex(out[0], out[1]), float(out[2]), float(out[3]), float(out[4])

    def overlap(self, other: "Simulator") -> Tuple[complex, float, float]:
        """(<self|other>, <self|self>, <other|other>) in FP64, on the device."""
        return self.compare(other)[:3]

    def fidelity(self, other: "Simulator") -> float:
        ov, na, nb = self.overlap(other)
        return abs(ov) ** 2 / (na * nb)

    def synchronize(self):
        _check(lib().kate_sim_synchronize(self._h), "synchronize")

    def measure_proba(self) -> np.ndarray:
        out = np.zeros(2 * self.n)
        _check(lib().kate_sim_measure_proba(self._h, _p(out, DP)), "measure")
        return out.reshape(self.n, 2)

    def expect_z(self, masks: Sequence[int]) -> np.ndarray:
        m = np.array(masks, dtype=np.uint64)
        out = np.zeros(len(m))
        _check(lib().kate_sim_expect_z(self._h, _p(m, U64P), len(m), _p(out, DP)), "expect_z")
        return out

    def norm(self) -> float:
        v = C.c_double()
        _check(lib().kate_sim_norm(self._h, C.byref(v)), "norm")
        return v.value

    def renormalize(self):
        _check(lib().kate_sim_renormalize(self._h), "renormalize")

    def state(self, real_order: bool = True) -> np.ndarray:
        out = np.zeros(1 << self.n, dtype=self.amp)
        _check(lib().kate_sim_read_state(self._h, out.ctypes.data, out.nbytes, 1 if real_order else 0), "read_state")
        return out

    def write_state(self, st: np.ndarray):
        a = np.ascontiguousarray(st, dtype=self.amp)
        _check(lib().kate_sim_write_state(self._h, a.ctypes.data, a.nbytes), "write_state")

    def device_state(self) -> int:
        p = VP()
        _check(lib().kate_sim_device_state(self._h, C.byref(p)), "device_state")
        return int(p.value or 0)

    def stream(self) -> int:
        p = VP()
        _check(lib().kate_sim_stream(self._h, C.byref(p)), "stream")
        return int(p.value or 0)

    def timings(self) -> dict:
        t = Timings()
        _check(lib().kate_sim_last_timings(self._h, C.byref(t)), "timings")
        return {f: getattr(t, f) for f, _ in Timings._fields_}

    def describe(self) -> str:
        buf = C.create_string_buffer(1 << 20)
        _check(lib().kate_sim_describe(self._h, buf, 1 << 20), "describe")
        return buf.value.decode()


class Batch:
    """Many circuits in one persistent launch (kate_batch_*): every state
    starts at |0...0>; a run returns <Z_q> for every qubit of every circuit."""

    def __init__(self, circuits: Sequence[Circuit], dtype: str = "c64", tile_qubits: int = 0, low_qubits: int = 0,
                 device: int = 0, mfma: int = 0, micro_bits: int = 0, euler: bool = True, direct: bool = True,
                 wgs_per_cu: int = 0, factor: bool = True):
        self.circuits = list(circuits)
        self.nq = [c.n for c in self.circuits]
        arr = (VP * len(self.circuits))(*[c._h for c in self.circuits])
        opt = Options({"c128": KATE_C128, "c64": KATE_C64}[dtype], tile_qubits, low_qubits, mfma, wgs_per_cu,
                      device, micro_bits, _plan_flags(euler, direct, factor=factor))
        h = VP()
        _check(lib().kate_batch_create_from_circuits(arr, len(self.circuits), C.byref(opt), C.byref(h)),
               "kate_batch_create")
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.kate_batch_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def capture(self):
        _check(lib().kate_batch_capture(self._h), "kate_batch_capture")

    def run(self) -> float:
        _check(lib().kate_batch_run(self._h), "kate_batch_run")
        ms = C.c_float()
        _check(lib().kate_batch_last_ms(self._h, C.byref(ms)), "kate_batch_last_ms")
        return float(ms.value)

    def info(self) -> dict:
        c, t, p, g = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        z = C.c_int64()
        _check(lib().kate_batch_info(self._h, C.byref(c), C.byref(z), C.byref(t), C.byref(p), C.byref(g)), "info")
        return {"circuits": c.value, "z_len": z.value, "tasks": t.value, "passes": p.value, "grid": g.value}

    def expect_z(self) -> Tuple[List[np.ndarray], np.ndarray]:
        """(<Z_q> array per circuit, norms)."""
        total = sum(self.nq)
        out = np.zeros(total)
        norms = np.zeros(len(self.nq))
        _check(lib().kate_batch_expect_z(self._h, _p(out, DP), total, _p(norms, DP)), "expect_z")
        res, o = [], 0
        for n in self.nq:
            res.append(out[o:o + n].copy())
            o += n
        return res, norms


class Server:
    """Continuous-batching request loop (kate_server_*): submit() is
    non-blocking; wait(id) returns the request's <Z_S> values and stats."""

    def __init__(self, dtype: str = "c128", streams: int = 4, max_active: int = 32, hbm_budget_bytes: int = 0,
                 tile_qubits: int = 0, low_qubits: int = 0, device: int = 0, mfma: int = 0, micro_bits: int = 0):
        opt = ServerOptions({"c128": KATE_C128, "c64": KATE_C64}[dtype], device, streams, max_active,
                            hbm_budget_bytes, tile_qubits, low_qubits, mfma, micro_bits)
        h = VP()
        _check(lib().kate_server_create(C.byref(opt), C.byref(h)), "kate_server_create")
        self._h = h
        self._nmasks = {}

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.kate_server_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, circuit: Circuit, masks: Sequence[int], theta_phi: Optional[np.ndarray] = None) -> int:
        m = np.array(masks, dtype=np.uint64)
        tp = None if theta_phi is None else np.ascontiguousarray(theta_phi, dtype=np.float64).ravel()
        rid = C.c_uint64()
        _check(lib().kate_server_submit_circuit(self._h, circuit._h, _p(tp, DP) if tp is not None else None,
                                                _p(m, U64P), len(m), C.byref(rid)), "kate_server_submit")
        self._nmasks[rid.value] = len(m)
        return rid.value

    def wait(self, rid: int) -> Tuple[np.ndarray, dict]:
        n = self._nmasks.pop(rid, 16)
        out = np.zeros(max(1, n))
        st = RequestStats()
        _check(lib().kate_server_wait(self._h, rid, _p(out, DP), n, C.byref(st)), "kate_server_wait")
        return out[:n], {f: getattr(st, f) for f, _ in RequestStats._fields_}

    def pending(self) -> dict:
        q, a = C.c_int32(), C.c_int32()
        c, f = C.c_int64(), C.c_int64()
        _check(lib().kate_server_pending(self._h, C.byref(q), C.byref(a), C.byref(c), C.byref(f)), "pending")
        return {"queued": q.value, "active": a.value, "completed": c.value, "largest_free_bytes": f.value}
