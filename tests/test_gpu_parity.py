"""GPU parity: libkate_hip.so on the MI355X against the oracle and Kate's own
outputs.  FP64 (complex128) must agree with Kate's CPU simulator within
1e-12 max-abs amplitude error; FP32 (complex64) within 1 - fidelity <= 1e-6."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import kate_amd as K
from fixture_io import unpack_gates
from kate_amd import workloads as W
from oracle import oracle as O
from util import compiled_from_circuit, compiled_from_fixture, fidelity, fixtures, load, to_real_order

pytestmark = pytest.mark.gpu

TOL64 = 1e-12   # max-abs amplitude error, complex128
FID32 = 1e-6    # 1 - fidelity, complex64

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)


def gpu_state(c, dtype="c128", proba_in=None, real_order=False, **opt):
    with K.Simulator(c, dtype=dtype, **opt) as sim:
        pr = sim.execute(proba_in)
        return sim.state(real_order=real_order), pr


@pytest.mark.parametrize("n,ngates,maxd,seed", [(1, 5, 1, 1), (3, 30, 3, 2), (7, 120, 3, 3), (12, 200, 4, 4),
                                                (14, 200, 4, 5), (16, 150, 3, 6), (13, 60, 6, 7)])
def test_random_circuits_match_oracle(gpu, n, ngates, maxd, seed):
    say("random", n, ngates, maxd, seed)
    gates = W.random_mixed(n, ngates, seed, max_dense=maxd)
    c = K.Circuit(n, gates)
    st, _ = gpu_state(c)
    ref = O.simulate(n, gates)
    assert np.max(np.abs(st - ref)) <= TOL64


@pytest.mark.parametrize("tile,low", [(10, 2), (11, 4), (12, 3), (13, 5), (12, 0)])
def test_tile_shapes_match_oracle(gpu, tile, low):
    say("tile", tile, low)
    n = 17
    gates = W.random_mixed(n, 160, 40 + tile, max_dense=3) + W.layered(n, 2, 9)
    st, _ = gpu_state(K.Circuit(n, gates), tile_qubits=tile, low_qubits=low)
    assert np.max(np.abs(st - O.simulate(n, gates))) <= TOL64


def test_config1_layered20_fp64(gpu):
    """Config 1: 20 qubits, 10 layers RX/RZ + CNOT ladder, vs Kate's CPU simulator."""
    say("config1")
    z, meta = load(fixtures("kate_config1_layered20.npz")[0])
    gates = unpack_gates(z, "in_")
    st, pr = gpu_state(K.Circuit(20, gates))
    err = np.max(np.abs(st[z["sample_index"]] - z["sample_amp"]))
    say("config1 max-abs vs Kate samples", err)
    assert err <= TOL64
    assert np.max(np.abs(st - O.simulate(20, gates))) <= TOL64
    assert np.max(np.abs(pr - z["proba"])) < 1e-9
    assert abs(np.vdot(st, st).real - z["norm"]) < 1e-12


SIM_FIXTURES = [p for p in fixtures() if "merge" not in p and "config1" not in p]


@pytest.mark.parametrize("path", SIM_FIXTURES, ids=lambda p: p.split("/")[-1][5:-4])
def test_kate_fixtures(gpu, path):
    z, meta = load(path)
    say("fixture", path)
    from test_passes import run_pipeline
    c = run_pipeline(K.Circuit(meta["n"], unpack_gates(z, "in_")), meta["pipeline"])
    if not meta["simulated"]:
        # Kate's compiler emitted swap instructions, which Kate's CPU
        # simulator ignores (the fixture holds no state): the program runs
        # against the oracle, whose swaps move amplitudes as Kate's GPU
        # simulator does
        prog = compiled_from_fixture(z, meta)
        st, _ = gpu_state(c)
        assert np.max(np.abs(st - O.run(prog))) <= TOL64
        real = to_real_order(st, prog.final_inverse_permutation)
        assert np.max(np.abs(real - O.simulate(meta["n"], unpack_gates(z, "in_")))) <= TOL64
        return
    pin = z["proba_in"] if "proba_in" in z else None
    st, pr = gpu_state(c, proba_in=pin)
    assert np.max(np.abs(st - z["state"])) <= TOL64
    assert np.max(np.abs(pr - z["proba"])) < 1e-9


@pytest.mark.parametrize("gpulog2", [1, 3])
def test_swap_programs(gpu, gpulog2):
    """Programs compiled for 2^g GPUs run on one: swap instructions become
    in-HBM remaps; real-order amplitudes equal the swap-free run's."""
    say("swap", gpulog2)
    n = 15
    gates = W.random_mixed(n, 150, 60 + gpulog2)
    c = K.Circuit(n, gates).schedule().group(6).allocate(gpulog2)
    assert any(k == 0 for k, _ in c.instructions())
    prog = compiled_from_circuit(c)
    st_virtual, _ = gpu_state(c)
    assert np.max(np.abs(st_virtual - O.run(prog))) <= TOL64
    st_real, _ = gpu_state(c, real_order=True)
    plain = K.Circuit(n, c.gates())  # not used for values; swap-free reference below
    del plain
    ref_real = to_real_order(O.run(prog), prog.final_inverse_permutation)
    assert np.max(np.abs(st_real - ref_real)) <= TOL64
    # the same circuit without allocation, scheduled identically
    c0 = K.Circuit(n, gates).schedule()
    st0, _ = gpu_state(c0)
    assert np.max(np.abs(st_real - st0)) <= 1e-12


def test_proba_expect_z_norm(gpu):
    say("observables")
    n = 18
    gates = W.brickwork(n, 12, 5, longrange_every=4, longrange_offset=7)
    c = K.Circuit(n, gates).schedule().group(7).allocate(0)  # virtual order != real order
    prog = compiled_from_circuit(c)
    ref_v = O.run(prog)
    ref_r = to_real_order(ref_v, prog.final_inverse_permutation)
    masks = [1 << q for q in range(n)] + [int(m) for m in W.zstring_masks(n, 8, 3)]
    with K.Simulator(c) as sim:
        pr = sim.execute()
        assert np.max(np.abs(pr - O.measure(n, ref_v, prog.final_inverse_permutation))) < 1e-9
        ez = sim.expect_z(masks)
        assert np.max(np.abs(ez - O.expect_z(n, ref_r, masks))) < 1e-12
        assert abs(sim.norm() - 1.0) < 1e-12
        st = sim.state(real_order=True)
        assert np.max(np.abs(st - ref_r)) <= TOL64


def test_product_state_inputs(gpu):
    say("product")
    n = 13
    rng = np.random.default_rng(3)
    tp = np.column_stack([rng.uniform(0, 1, n), rng.uniform(0, 6, n)])
    gates = W.random_mixed(n, 100, 77)
    c = K.Circuit(n, gates).allocate(0)
    prog = compiled_from_circuit(c)
    st, pr = gpu_state(c, proba_in=tp)
    ref = O.run(prog, tp)
    assert np.max(np.abs(st - ref)) <= TOL64
    assert np.max(np.abs(pr - O.measure(n, ref, prog.final_inverse_permutation))) < 1e-9


def test_qft_basis_closed_form(gpu):
    say("qft")
    n, x = 16, 0b1011001110001011
    st, _ = gpu_state(K.Circuit(n, W.qft(n)), proba_in=W.basis_input(n, x))
    rev = int(format(x, f"0{n}b")[::-1], 2)
    expect = np.exp(2j * np.pi * rev * np.arange(1 << n) / (1 << n)) / 2 ** (n / 2)
    assert np.max(np.abs(st - expect)) < 1e-12


def test_fused_programs_fp64(gpu):
    say("fused64")
    n = 16
    gates = W.layered(n, 6, 4)
    c = K.Circuit(n, gates).schedule().fuse(5).group(10)
    st, _ = gpu_state(c)
    assert np.max(np.abs(st - O.run(compiled_from_circuit(c)))) <= TOL64


@pytest.mark.parametrize("kind", ["layered", "brickwork", "mixed"])
def test_split_fused_gates_match_oracle(gpu, kind):
    """Fused programs whose dense 2-qubit gates the planner applies as
    Kronecker factors and a layout CNOT: the state equals the oracle's run
    of the same fused program (FP64, 1e-12) and the run with the split off
    to rounding."""
    say("split", kind)
    n = 18
    gates = {"layered": W.layered(n, 8, 7), "brickwork": W.brickwork(n, 6, 8, longrange_offset=9),
             "mixed": W.random_mixed(n, 120, 9, max_dense=2) + W.layered(n, 3, 9)}[kind]
    c = K.Circuit(n, gates).schedule().fuse(5)
    ref = O.run(compiled_from_circuit(c))
    with K.Simulator(c, dtype="c128") as a, K.Simulator(K.Circuit(n, gates).schedule().fuse(5), dtype="c128",
                                                        split=False) as b:
        a.execute()
        b.execute()
        sa, sb = a.state(real_order=False), b.state(real_order=False)
    assert np.max(np.abs(sa - ref)) <= TOL64
    assert np.max(np.abs(sa - sb)) <= 1e-13


@pytest.mark.parametrize("n,fuse", [(20, 0), (20, 5), (16, 4)])
def test_fp32_fidelity(gpu, n, fuse):
    say("fp32", n, fuse)
    gates = W.layered(n, 10, 1)
    c = K.Circuit(n, gates)
    if fuse:
        c.schedule().fuse(fuse)
    st32, _ = gpu_state(c, dtype="c64")
    ref = O.simulate(n, gates)
    assert st32.dtype == np.complex64
    assert 1 - fidelity(st32, ref) <= FID32
    assert np.max(np.abs(st32 - ref)) < 1e-5


def test_mirror_circuit_24q(gpu):
    """Size-independent property at a larger size: U then U^dagger returns |0>."""
    say("mirror")
    n = 24
    gates = W.brickwork(n, 8, 12, longrange_every=4, longrange_offset=16)
    inv = [W.Gate(0, g.qubits, 0, g.mat.conj().T.copy()) for g in reversed(gates)]
    for dtype, tol in (("c128", 1e-12), ("c64", 1e-5)):
        with K.Simulator(K.Circuit(n, gates + inv), dtype=dtype) as sim:
            sim.execute()
            ez = sim.expect_z([1 << q for q in range(n)])
            assert np.max(np.abs(ez - 1.0)) < 10 * tol
            st = sim.state(real_order=False)
            assert abs(st[0] - 1) < tol and np.max(np.abs(st[1:])) < tol


def test_repeatable_bitwise(gpu):
    say("repeat")
    n = 19
    gates = W.random_mixed(n, 120, 5)
    with K.Simulator(K.Circuit(n, gates)) as sim:
        a = sim.execute()
        s1 = sim.state()
        b = sim.execute()
        s2 = sim.state()
    assert np.array_equal(a, b) and np.array_equal(s1, s2)


def test_drop_in_binary(gpu, tmp_path):
    """The unchanged Kate user program, linked against the library, agrees with the oracle."""
    say("drop-in")
    exe = tmp_path / "drop_in"
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-I",
           os.path.join(ROOT, "kate_amd", "kate"), os.path.join(ROOT, "tests", "cpp", "drop_in_main.cpp"), "-o",
           str(exe), "-L", os.path.join(ROOT, "kate_amd"), "-lkate_hip", "-Wl,-rpath," + os.path.join(ROOT, "kate_amd")]
    subprocess.run(cmd, check=True)
    n = 11
    out = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-1000:] + out.stderr[-1000:]
    amps = np.zeros(1 << n, dtype=np.complex128)
    for line in out.stdout.splitlines():
        if line.startswith("A "):
            _, i, re_, im_ = line.split()
            amps[int(i)] = float(re_) + 1j * float(im_)
    gates = W.qft(n) + [W.Gate(W.H, (0,)), W.Gate(W.CNOT, (1, 0)),
                        W.Gate(0, (n - 1,), 0, np.array([[0, 1], [1, 0]], dtype=np.complex128))]
    tp = np.array([[float(i % 2), 0.0] for i in range(n)])
    ref = O.simulate(n, gates, tp)
    assert np.max(np.abs(amps - ref)) < 1e-12


@pytest.mark.parametrize("mode", [1, 3], ids=["fp32", "bf16x3"])
@pytest.mark.parametrize("k", [4, 5])
def test_mfma_dense_gates_fp32(gpu, k, mode):
    """Dense 4- and 5-qubit gates on the matrix cores (complex64, FP32 accumulate)."""
    say("mfma dense", k)
    n = 16
    rng = np.random.default_rng(50 + k)
    gates = []
    for i in range(12):
        qs = tuple(int(x) for x in rng.choice(n, size=k, replace=False))
        gates.append(W.Gate(0, qs, 0, W.haar_unitary(rng, 1 << k)))
        gates.append(W.Gate(W.H, (qs[1],)))
    ref = O.simulate(n, gates)
    st, _ = gpu_state(K.Circuit(n, gates), dtype="c64", mfma=mode)
    say("1-F", 1 - fidelity(st, ref), "max-abs", np.max(np.abs(st - ref)))
    assert 1 - fidelity(st, ref) <= FID32
    assert np.max(np.abs(st - ref)) < 1e-5


@pytest.mark.parametrize("mode", [2, 3], ids=["fp32", "bf16x3"])
@pytest.mark.parametrize("tile", [12, 13])
def test_mfma_window_fusion_fp32(gpu, tile, mode):
    """Every 5-qubit window of gates composed in FP64 and applied on the matrix cores."""
    say("mfma windows", tile)
    n = 18
    gates = W.layered(n, 6, 2) + W.random_mixed(n, 80, 9, max_dense=3)
    ref = O.simulate(n, gates)
    st, _ = gpu_state(K.Circuit(n, gates), dtype="c64", mfma=mode, tile_qubits=tile)
    say("1-F", 1 - fidelity(st, ref), "max-abs", np.max(np.abs(st - ref)))
    assert 1 - fidelity(st, ref) <= FID32
    assert np.max(np.abs(st - ref)) < 1e-5


def test_overlap_fidelity_on_device(gpu):
    say("overlap")
    n = 20
    gates = W.layered(n, 6, 3)
    with K.Simulator(K.Circuit(n, gates), dtype="c128") as a, K.Simulator(K.Circuit(n, gates), dtype="c64") as b:
        a.execute()
        b.execute()
        ov, na, nb, mx = a.compare(b)
        sa, sb = a.state(real_order=False), b.state(real_order=False).astype(np.complex128)
        assert abs(mx - np.abs(sa - sb).max()) <= 1e-15
        assert abs(ov - np.vdot(sa, sb)) < 1e-10
        assert abs(na - np.vdot(sa, sa).real) < 1e-12 and abs(nb - np.vdot(sb, sb).real) < 1e-10
        assert 1 - a.fidelity(b) <= FID32


@pytest.mark.parametrize("n,depth,seed,dtype,mb,direct,tile", [
    (18, 6, 11, "c128", 4, True, 12), (18, 6, 11, "c128", 4, False, 12), (16, 5, 12, "c128", 3, True, 12),
    (20, 8, 13, "c128", 4, True, 10), (20, 8, 13, "c64", 4, True, 12), (17, 6, 14, "c64", 4, False, 11)])
def test_factored_euler_steps_match_oracle(gpu, n, depth, seed, dtype, mb, direct, tile):
    """Light passes apply a step's lone Euler-form op in the scaled form
    (fkron), whose stores undo the exchanged outputs of its flip mask: the
    state matches the oracle (complex128 within 1e-12, complex64 within
    1 - F <= 1e-6) and the unscaled plan to rounding, through LDS stores,
    direct stores and 3-bit micro-steps."""
    say("factored", n, depth, seed, dtype, mb, direct, tile)
    gates = W.layered(n, depth, seed)
    c = K.Circuit(n, gates)
    _, _, text = c.plan(dtype, tile_qubits=tile, micro_bits=mb, direct=direct)
    assert "light" in text
    opt = dict(tile_qubits=tile, micro_bits=mb, direct=direct)
    st, _ = gpu_state(c, dtype, **opt)
    st0, _ = gpu_state(c, dtype, factor=False, **opt)
    ref = O.simulate(n, gates)
    if dtype == "c128":
        assert np.max(np.abs(st - ref)) <= TOL64
        assert np.max(np.abs(st - st0)) <= 1e-13
    else:
        assert 1.0 - fidelity(st, ref) <= FID32
        assert 1.0 - fidelity(st, st0) <= FID32


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_factored_single_step_direct_pass(gpu, dtype, monkeypatch):
    """A pass of one micro-step, loaded from HBM into registers and stored
    straight back (direct first and last step), whose lone op is a scaled
    Euler-form op with a non-zero flip mask: from |0...0> and from a product
    state, against the oracle."""
    n = 13
    gates = [g for g in W.layered(n, 1, 1) if len(g.qubits) == 1 and g.qubits[0] >= 9]
    monkeypatch.setenv("KATE_DESCRIBE_OPS", "1")
    _, _, text = K.Circuit(n, gates).plan(dtype)
    monkeypatch.delenv("KATE_DESCRIBE_OPS")
    m = [int(x) for x in re.findall(r" fkron:(\d+)", text)]
    assert "micro=1 " in text and "direct=first,last" in text and m and m[0] >> 4
    rng = np.random.default_rng(5)
    tp = np.column_stack([rng.uniform(0, 1, n), rng.uniform(0, 6, n)])
    for proba_in in (None, tp):
        st, _ = gpu_state(K.Circuit(n, gates), dtype, proba_in=proba_in)
        ref = O.simulate(n, gates, proba_in)
        if dtype == "c128":
            assert np.max(np.abs(st - ref)) <= TOL64
        else:
            assert 1.0 - fidelity(st, ref) <= FID32


def test_factored_steps_special_angles(gpu):
    """Scaled Euler steps at the edges of their case split: RX angles whose
    cos and sin halves tie (pi/2), vanish (0, pi: the Euler form then falls
    back to the plain product) or nearly vanish (1e-9, pi - 1e-9, so that
    t = y/x or x/y is tiny), beside random ones, in the layered structure;
    against the oracle and the unscaled plan."""
    say("factored special angles")
    n = 16
    rng = np.random.default_rng(31)
    special = [0.0, np.pi / 2, np.pi, np.pi - 1e-9, 1e-9, 3 * np.pi / 2, 2 * np.pi - 1e-7, -np.pi / 2]
    gates = []
    for layer in range(8):
        for q in range(n):
            th = special[(q + layer) % len(special)] if (q + layer) % 3 else float(rng.uniform(0, 2 * np.pi))
            gates.append(W.Gate(W.DENSE, (q,), 0, W.rx(th)))
            gates.append(W.Gate(W.DENSE, (q,), 0, W.rz(float(rng.uniform(0, 2 * np.pi)))))
        for i in range(n - 1):
            gates.append(W.Gate(W.CNOT, (i + 1, i)))
    ref = O.simulate(n, gates)
    for dtype in ("c128", "c64"):
        st, _ = gpu_state(K.Circuit(n, gates), dtype)
        st0, _ = gpu_state(K.Circuit(n, gates), dtype, factor=False)
        if dtype == "c128":
            assert np.max(np.abs(st - ref)) <= TOL64
            assert np.max(np.abs(st - st0)) <= 1e-13
        else:
            assert 1.0 - fidelity(st, ref) <= FID32


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_factored_steps_in_swap_programs(gpu, dtype):
    """A layered circuit compiled for 4 GPUs: its gate passes run the scaled
    Euler steps on the light kernel, its swaps run as remap passes; against
    the oracle run of the compiled program and, in real qubit order, against
    the swap-free circuit."""
    say("factored swap program", dtype)
    n = 17
    gates = W.layered(n, 6, 77)
    c = K.Circuit(n, gates).schedule().group(8).allocate(2)
    assert any(k == 0 for k, _ in c.instructions())
    _, remaps, text = c.plan(dtype)
    assert remaps > 0 and " light" in text
    prog = compiled_from_circuit(c)
    st, _ = gpu_state(c, dtype)
    ref = O.run(prog)
    if dtype == "c128":
        assert np.max(np.abs(st - ref)) <= TOL64
    else:
        assert 1.0 - fidelity(st, ref) <= FID32
    st_real, _ = gpu_state(c, dtype, real_order=True)
    st0, _ = gpu_state(K.Circuit(n, gates), dtype)
    if dtype == "c128":
        assert np.max(np.abs(st_real - st0)) <= 1e-12
    else:
        assert 1.0 - fidelity(st_real, st0) <= FID32


def dense_gates(n, count, k, seed):
    rng = np.random.default_rng(seed)
    return [W.Gate(W.DENSE, tuple(int(x) for x in rng.choice(n, size=k, replace=False)), 0,
                   W.haar_unitary(rng, 1 << k)) for _ in range(count)]


@pytest.mark.parametrize("dtype,tile,mb,matrix5", [("c128", 12, 4, True), ("c128", 10, 4, True), ("c128", 11, 3, True),
                                                   ("c128", 12, 3, True), ("c64", 12, 4, False), ("c64", 13, 4, False),
                                                   ("c64", 10, 4, False), ("c64", 12, 4, True), ("c64", 13, 4, True),
                                                   ("c64", 10, 4, True)])
def test_dense_five_qubit_gates(gpu, dtype, tile, mb, matrix5):
    """Dense 5-qubit gates against the oracle, on every tile width and both
    register widths, with 1- to 4-qubit gates between them: register-blocked
    LDS steps (each thread holds a group of 32 amplitudes, the matrix rows
    are wave-uniform), and for complex64 by default the FP32 matrix cores."""
    say("dense5", dtype, tile, mb, matrix5)
    n = 16
    gates = dense_gates(n, 24, 5, 70 + tile) + W.random_mixed(n, 60, 71, max_dense=4) + dense_gates(n, 12, 5, 72)
    st, _ = gpu_state(K.Circuit(n, gates), dtype=dtype, tile_qubits=tile, micro_bits=mb, matrix5=matrix5)
    ref = O.simulate(n, gates)
    if dtype == "c128":
        assert np.max(np.abs(st - ref)) <= TOL64
    else:
        assert 1.0 - abs(np.vdot(ref, st.astype(np.complex128))) ** 2 <= 1e-6
