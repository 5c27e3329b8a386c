"""Parity at the full workload sizes (30-31 qubits, 16-32 GiB states) through
size-independent properties, reduced on the device (no 16 GiB host copies):

* mirror: the 30-qubit layered circuit followed by its inverse returns
  |0...0> amplitude by amplitude (max-abs against a |0...0> state);
* precision: the headline circuit in complex64 against complex128,
  1 - fidelity <= 1e-6;
* swaps: the 31-qubit brickwork compiled for 8 GPUs by Kate's allocate(3)
  (its swaps run as in-HBM remap passes) gives the same <Z_S> as the plain
  program; the norm stays 1.
"""
import numpy as np
import pytest

import kate_amd as K
from kate_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def inverse(gates):
    out = []
    for g in reversed(gates):
        if g.id == W.DENSE:
            out.append(W.Gate(W.DENSE, g.qubits, 0, g.mat.conj().T.copy()))
        elif g.id == W.CNOT:
            out.append(g)
        else:
            raise ValueError("inverse of this gate kind is not needed here")
    return out


def test_layered_30q_mirror(gpu):
    n = 30
    gates = W.layered(n, 10, 1)
    circ = gates + inverse(gates)
    say("fullsize mirror", n, len(circ))
    with K.Simulator(K.Circuit(n, []), dtype="c128") as zero:
        zero.init_basis(0)
        for dtype, tol in (("c128", 1e-11), ("c64", 5e-5)):
            with K.Simulator(K.Circuit(n, circ), dtype=dtype) as sim:
                sim.init_basis(0)
                sim.run()
                ov, na, nb, mx = zero.compare(sim)
                say(dtype, "max-abs vs |0>", mx, "norm", nb)
                assert mx < tol
                assert abs(nb - 1.0) < (1e-12 if dtype == "c128" else 1e-5)


def test_headline_fp32_fidelity(gpu):
    n = 30
    gates = W.layered(n, 20, 1)
    say("fullsize fp32 fidelity")
    with K.Simulator(K.Circuit(n, gates), dtype="c128") as a, K.Simulator(K.Circuit(n, gates), dtype="c64") as b:
        a.init_basis(0)
        a.run()
        b.init_basis(0)
        b.run()
        ov, na, nb, mx = a.compare(b)
        infid = 1.0 - abs(ov) ** 2 / (na * nb)
        say("1-F", infid, "max-abs", mx)
        assert abs(na - 1.0) < 1e-12
        assert infid <= 1e-6
        assert mx < 1e-6


def test_brickwork_31q_swap_plan_matches_plain(gpu):
    n = 31
    gates = W.brickwork(n, 24, 5)
    masks = [1 << q for q in range(n)] + [int(m) for m in W.zstring_masks(n, 8, 3)]
    say("fullsize 31q swap plan")
    with K.Simulator(K.Circuit(n, gates), dtype="c128") as sim:
        sim.init_basis(0)
        sim.run()
        z_plain = sim.expect_z(masks)
        assert abs(sim.norm() - 1.0) < 1e-12
    c = K.Circuit(n, gates).schedule().group(8).allocate(3)
    assert any(k == 0 for k, _ in c.instructions())
    with K.Simulator(c, dtype="c128") as sim:
        sim.init_basis(0)
        sim.run()
        assert sim.timings()["remap_passes"] > 0
        z_swap = sim.expect_z(masks)
        assert abs(sim.norm() - 1.0) < 1e-12
    say("max |<Z> plain - <Z> swap plan|", np.max(np.abs(z_plain - z_swap)))
    assert np.max(np.abs(z_plain - z_swap)) < 1e-12


def test_c2_28q_complex64_fused_forms(gpu):
    """The 28-qubit depth-20 layered circuit: complex64 plain, and
    GateMerger(5)-fused on the register path, on FP32 matrix-core windows
    and on split-bf16 matrix-core windows, each against the complex128
    state on the device: 1 - fidelity <= 1e-6 and max-abs amplitude error
    (printed, bounded)."""
    n = 28
    gates = W.layered(n, 20, 2)
    say("C2 28q complex64 forms")
    with K.Simulator(K.Circuit(n, gates), dtype="c128") as ref:
        ref.execute()
        for label, fuse, mfma in (("plain", False, 0), ("fused valu", True, 0), ("fused mfma fp32", True, 2),
                                  ("fused mfma bf16x3", True, 3)):
            c = K.Circuit(n, gates).schedule().fuse(5) if fuse else K.Circuit(n, gates)
            with K.Simulator(c, dtype="c64", mfma=mfma) as s32:
                s32.execute()
                ov, na, nb, mx = ref.compare(s32)
                infid = 1.0 - abs(ov) ** 2 / (na * nb)
                say(f"C2 {label}: 1-F {infid:.3e} max-abs {mx:.3e} norm {nb:.9f}")
                assert abs(na - 1.0) < 1e-12
                assert infid <= 1e-6, label
                assert mx < 1e-6, label


def test_c3_32q_complex64_brickwork(gpu):
    """The 32-qubit complex64 brickwork (depth 24, long-range pairs every
    4th layer, 32 GiB state): the plain program, and Kate's allocate(3)
    swap plan (in-HBM remap passes), against the same circuit in complex128
    (64 GiB) on the device: <Z_S> for every qubit and 8 Z strings, the
    fidelity and the norm."""
    n = 32
    gates = W.brickwork(n, 24, 5)
    masks = [1 << q for q in range(n)] + [int(m) for m in W.zstring_masks(n, 8, 3)]
    say("C3 32q complex128 reference")
    with K.Simulator(K.Circuit(n, gates), dtype="c128") as ref:
        ref.execute()
        z_ref = ref.expect_z(masks)
        assert abs(ref.norm() - 1.0) < 1e-12
        c_swap = K.Circuit(n, gates).schedule().group(8).allocate(3)
        assert any(k == 0 for k, _ in c_swap.instructions())
        for label, c in (("plain", K.Circuit(n, gates)), ("allocate(3) swap plan", c_swap)):
            with K.Simulator(c, dtype="c64") as s32:
                s32.execute()
                if label != "plain":
                    assert s32.timings()["remap_passes"] > 0
                z = s32.expect_z(masks)
                nrm = s32.norm()
                zerr = float(np.max(np.abs(z - z_ref)))
                msg = f"C3 32q complex64 {label}: max <Z_S> error {zerr:.3e} norm {nrm:.9f}"
                if label == "plain":
                    ov, na, nb, mx = ref.compare(s32)
                    infid = 1.0 - abs(ov) ** 2 / (na * nb)
                    msg += f" 1-F {infid:.3e} max-abs {mx:.3e}"
                    assert infid <= 1e-6
                say(msg)
                assert zerr < 1e-5, label
                assert abs(nrm - 1.0) < 1e-5, label
