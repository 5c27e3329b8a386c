"""The C ABI surface: every declared symbol is exported, arguments and
programs are validated, and nothing silently falls back to the CPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import kate_amd as K
from kate_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "kate_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(kate_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    syms = declared_symbols()
    assert len(syms) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", K.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (kate_\w+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    assert sorted(K.EXPORTS) == syms


def test_library_loads_and_reports_devices():
    n = K.device_count()
    assert n >= 0
    if n == 0:
        with pytest.raises(K.KateError, match="no GPU"):
            K.Simulator(K.Circuit(3, [W.Gate(W.H, (0,))]))


def test_circuit_argument_errors():
    c = K.Circuit(4)
    with pytest.raises(K.KateError, match="out of range"):
        c.append(W.Gate(W.H, (4,)))
    with pytest.raises(K.KateError, match="present twice"):
        K.lib().kate_circuit_append  # symbol exists
        qs = np.array([1, 1], dtype=np.int32)
        rc = K.lib().kate_circuit_append(c._h, W.CNOT, 2, qs.ctypes.data_as(K.I32P), 0, None)
        K._check(rc, "append")
    with pytest.raises(K.KateError, match="gate id"):
        qs = np.array([1], dtype=np.int32)
        K._check(K.lib().kate_circuit_append(c._h, 6, 1, qs.ctypes.data_as(K.I32P), 0, None), "append RX")
    assert c.info()[1] == 0


def _program(nq, recs, pool=None, **kw):
    arr = (K._GateRec * max(1, len(recs)))()
    for i, (gid, qs, oa, moff) in enumerate(recs):
        arr[i].id, arr[i].nq, arr[i].optarg, arr[i].matrix_offset = gid, len(qs), oa, moff
        for j, q in enumerate(qs):
            arr[i].qubits[j] = q
    pool = np.zeros(2, dtype=np.float64) if pool is None else pool

    class Prog(C.Structure):
        _fields_ = [("nqubits", C.c_int32), ("ngates", C.c_int32), ("gates", C.c_void_p),
                    ("matrix_pool", C.c_void_p), ("matrix_pool_len", C.c_int64), ("ngroups", C.c_int32),
                    ("group_end", C.c_void_p), ("group_mask", C.c_void_p), ("ninstr", C.c_int32),
                    ("instr_kind", C.c_void_p), ("instr_npairs", C.c_void_p), ("instr_pair_offset", C.c_void_p),
                    ("pairs", C.c_void_p), ("initial_permutation", C.c_void_p),
                    ("final_inverse_permutation", C.c_void_p)]
    p = Prog(nq, len(recs), C.addressof(arr), pool.ctypes.data, pool.size // 2)
    for k, v in kw.items():
        setattr(p, k, v)
    return p, (arr, pool)


def plan(p):
    passes, remaps = C.c_int32(), C.c_int32()
    buf = C.create_string_buffer(1 << 16)
    rc = K.lib().kate_plan_describe(C.byref(p), None, C.byref(passes), C.byref(remaps), buf, 1 << 16)
    return rc, K.lib().kate_last_error().decode()


def test_program_validation():
    ok, keep = _program(3, [(2, [0], 0, -1), (3, [1, 0], 0, -1)])
    assert plan(ok)[0] == 0
    bad, keep = _program(3, [(2, [3], 0, -1)])
    assert plan(bad) == (-5, "gate 0: qubit out of range")
    bad, keep = _program(3, [(0, [0], 0, 5)], pool=np.zeros(8))
    assert plan(bad)[0] == -5 and "matrix outside the pool" in plan(bad)[1]
    bad, keep = _program(3, [(7, [0], 0, -1)])
    assert "unsupported gate id 7" in plan(bad)[1]
    bad, keep = _program(3, [(3, [0, 1, 2], 0, -1)])
    assert "CNOT acts on exactly 2" in plan(bad)[1]
    perm = np.array([0, 0, 1], dtype=np.int32)
    bad, keep = _program(3, [(2, [0], 0, -1)], final_inverse_permutation=perm.ctypes.data)
    assert "not a permutation" in plan(bad)[1]
    ends = np.array([1], dtype=np.int32)
    bad, keep = _program(3, [(2, [0], 0, -1), (2, [1], 0, -1)], ngroups=1, group_end=ends.ctypes.data)
    assert "do not cover" in plan(bad)[1]


def test_null_handles_are_rejected():
    assert K.lib().kate_sim_run(None) == -1
    assert K.lib().kate_sim_destroy(None) == 0
    assert K.lib().kate_circuit_schedule(None) == -1


def test_qcx_export_format(tmp_path):
    c = K.Circuit(3, [W.Gate(W.H, (0,)), W.Gate(W.CNOT, (1, 0)), W.Gate(W.CRK, (2, 0), 3),
                      W.Gate(W.TOFFOLI, (0, 1, 2)), W.Gate(0, (1, 2), 0, np.eye(4, dtype=complex))])
    path = tmp_path / "c.qcx"
    c.export_qcx(str(path))
    # Gate::transformqcx / Circuit::exportqcx (basic_host_types.hpp:270-295, Circuit.hpp:763-774)
    assert path.read_text() == "QUBITS 3\nH 0 \nCNOT 1 0 \nU 2 0 3\nTOFFOLI 0 1 2 \nXX 1 2 \nBEGIN MEASUREMENT"


def test_drop_in_header_compiles_and_links(tmp_path):
    """User code written against Kate's API compiles unchanged against the
    drop-in headers and links against libkate_hip.so."""
    exe = tmp_path / "drop_in"
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-I",
           os.path.join(ROOT, "kate_amd", "kate"), os.path.join(ROOT, "tests", "cpp", "drop_in_main.cpp"), "-o",
           str(exe), "-L", os.path.join(ROOT, "kate_amd"), "-lkate_hip", "-Wl,-rpath," + os.path.join(ROOT, "kate_amd")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
