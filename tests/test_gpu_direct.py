"""Direct first / last micro-steps (the first register step of a pass reads
its amplitudes from HBM, the last one writes them back, without the LDS fill
and drain): the same arithmetic as the plain pass, so states agree bit for
bit with direct=False, and with the oracle to the usual tolerance.  Covers
both register widths, every tile shape, complex64, lazy-CNOT layouts (the
CNOT ladders), swap programs (remap passes) and batches."""
import numpy as np
import pytest

import kate_amd as K
from kate_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def state_of(c, dtype="c128", **opt):
    with K.Simulator(c, dtype=dtype, **opt) as sim:
        sim.execute()
        return sim.state()


CASES = [
    ("layered", 14, lambda n: W.layered(n, 6, 21)),
    ("mixed", 13, lambda n: W.random_mixed(n, 200, 22, max_dense=4) + W.layered(n, 2, 23)),
    ("brick", 16, lambda n: W.brickwork(n, 8, 24, longrange_offset=8)),
    ("haar", 15, lambda n: W.haar_layers(n, 4, 25)),
    ("qft", 14, lambda n: W.qft(n)),
]


@pytest.mark.parametrize("dtype,mb", [("c128", 3), ("c128", 4), ("c64", 4)])
@pytest.mark.parametrize("tile", [10, 12, 13])
@pytest.mark.parametrize("name,n,gen", CASES, ids=[c[0] for c in CASES])
def test_direct_steps_bit_exact(gpu, dtype, mb, tile, name, n, gen):
    say("direct", dtype, mb, tile, name, n)
    gates = gen(n)
    _, _, text = K.Circuit(n, gates).plan(dtype, tile_qubits=tile, micro_bits=mb)
    a = state_of(K.Circuit(n, gates), dtype, tile_qubits=tile, micro_bits=mb)
    b = state_of(K.Circuit(n, gates), dtype, tile_qubits=tile, micro_bits=mb, direct=False)
    assert np.array_equal(a, b), (name, "direct=" in text)
    tol = 1e-12 if dtype == "c128" else 2e-6
    assert np.max(np.abs(a - O.simulate(n, gates))) <= tol


def test_direct_swap_program_and_20_qubits(gpu):
    say("direct 20q swap program")
    n = 20
    gates = W.brickwork(n, 10, 27, longrange_offset=10)
    for mk in (lambda: K.Circuit(n, gates), lambda: K.Circuit(n, gates).schedule().group(8).allocate(3)):
        a = state_of(mk())
        b = state_of(mk(), direct=False)
        assert np.array_equal(a, b)


def test_direct_single_step_pass_at_28_qubits(gpu):
    """The 28-qubit layered plan with 3 low qubits has a pass of one
    micro-step, loaded and stored directly: the store of one wave must not
    overtake a load of another wave of the tile (the final layout moves
    amplitudes between threads).  The circuit followed by its inverse, whose
    plan holds the same single-step pass, must return |0...0> amplitude by
    amplitude (compared on the device against a |0...0> state)."""
    from test_gpu_fullsize import inverse

    say("direct 28q single-step pass (mirror)")
    n = 28
    gates = W.layered(n, 20, 2)
    circ = gates + inverse(gates)
    _, _, text = K.Circuit(n, circ).plan("c128", low_qubits=3, reorder=False)
    assert any(" micro=1 " in ln and "direct=first,last" in ln for ln in text.splitlines())
    with K.Simulator(K.Circuit(n, []), dtype="c128") as zero, \
            K.Simulator(K.Circuit(n, circ), dtype="c128", low_qubits=3, reorder=False) as a:
        zero.init_basis(0)
        a.execute()
        _, na, nb, mx = zero.compare(a)
    say("max-abs vs |0>", mx)
    assert mx <= 1e-11 and abs(nb - 1.0) <= 1e-12


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_execute_makes_the_basis_state_in_its_first_pass(gpu, dtype):
    """execute() from |0...0> lets the first pass make its tiles in registers
    instead of running the init kernel, and follows the support of the state
    (tiles of qubits no gate has touched yet are neither read nor launched
    until a pass needs them; the last pass writes every tile): the same
    state, bit for bit, as init_basis(0) followed by run(), including
    circuits that never touch some qubits and CNOTs whose controls are still
    |0>."""
    say("fused init", dtype)
    partial = W.random_mixed(10, 150, 34, max_dense=3)  # qubits 10..17 never touched
    ctrl = [W.Gate(W.CNOT, (q, q + 8)) for q in range(8)] + W.layered(16, 2, 35)  # controls still |0>
    for n, gates in ((14, W.random_mixed(14, 120, 31, max_dense=4)), (20, W.layered(20, 6, 32)),
                     (18, W.brickwork(18, 6, 33, longrange_offset=9)), (18, partial), (16, ctrl),
                     (16, W.qft(16)), (17, W.haar_layers(17, 3, 36))):
        for mk in (lambda: K.Circuit(n, gates), lambda: K.Circuit(n, gates).schedule().group(8).allocate(2)):
            with K.Simulator(mk(), dtype=dtype) as sim:
                sim.execute()
                a = sim.state()
                sim.init_basis(0)
                sim.run()
                b = sim.state()
            assert np.array_equal(a, b)


def test_support_skipping_at_30_qubits(gpu):
    """The 30-qubit layered circuit from |0...0>: execute() with the support
    skipping against execute() without it, on the device."""
    say("support 30q")
    n = 30
    gates = W.layered(n, 20, 1)
    with K.Simulator(K.Circuit(n, gates), dtype="c128") as a, \
            K.Simulator(K.Circuit(n, gates), dtype="c128", support=False) as b:
        a.execute()
        b.execute()
        _, na, nb, mx = a.compare(b)
    assert mx == 0.0 and abs(na - 1.0) <= 1e-12


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_execute_measures_through_trailing_permutations(gpu, dtype):
    """A circuit that ends with CNOT layers leaves its last passes as pure
    permutations of the state: execute() measures through them (the sums
    read amplitude i where it still is) and runs them when the state is
    next read.  The proba_state must equal the one measured on the finished
    state, and the finished state the one of a run without the deferral."""
    say("deferred trailing passes", dtype)
    for n, gates in ((20, W.layered(20, 3, 41)), (22, W.layered(22, 3, 42)),
                     (24, W.haar_layers(24, 1, 43) + W.layered(24, 2, 44))):
        # 12-qubit tiles: the plans of these circuits end with permutation
        # passes (13-qubit complex64 tiles fit the 20-qubit one in 2 passes);
        # without the scatter store, which folds them into the last gate pass
        _, _, text = K.Circuit(n, gates).plan(dtype, tile_qubits=12, scatter=False)
        assert " micro=0 " in text.splitlines()[-1] + " "  # the last pass only permutes
        with K.Simulator(K.Circuit(n, gates), dtype=dtype, defer=True, tile_qubits=12, scatter=False) as a, \
                K.Simulator(K.Circuit(n, gates), dtype=dtype, defer=False, tile_qubits=12, scatter=False) as b:
            pa = a.execute()
            pb = b.execute()
            sa = a.state()  # runs the deferred passes
            pa2 = a.measure_proba()
            sb = b.state()
        tol = 1e-12 if dtype == "c128" else 1e-5
        assert np.max(np.abs(pa - pb)) <= tol and np.max(np.abs(pa - pa2)) <= tol
        assert np.array_equal(sa, sb)
        if dtype == "c128" and n <= 20:
            ref = O.simulate(n, gates)
            assert np.max(np.abs(sa - ref)) <= 1e-12
            assert np.max(np.abs(pa - O.measure(n, ref))) <= 1e-9


@pytest.mark.parametrize("seed", range(6))
def test_runtime_savings_match_the_plain_run(gpu, seed):
    """Random programs (all gate kinds, some compiled with Kate's swap
    allocator): execute() with the direct steps, the support skipping and
    the deferred trailing passes gives the same proba_state and the same
    state as with all three switched off: the support skipping and the
    deferral bit for bit; the direct first / last steps compile the same
    register ops a second time, whose FMA contraction can differ in the last
    bit (observed 2e-18 on amplitudes, 2e-15 on proba angles)."""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(13, 21))
    kind = seed % 3
    if kind == 0:
        gates = W.random_mixed(n, 150, 200 + seed, max_dense=4) + W.layered(n, 2, seed)
    elif kind == 1:
        gates = W.brickwork(n, 6, 300 + seed, longrange_offset=n // 2) + [W.Gate(W.CNOT, (q + 1, q)) for q in range(n - 1)]
    else:
        gates = W.layered(n, 4, 400 + seed)
    say("savings vs plain", seed, n, kind)
    for alloc in (0, 2):
        def mk():
            c = K.Circuit(n, gates)
            return c.schedule().group(8).allocate(alloc) if alloc else c
        with K.Simulator(mk(), dtype="c128", defer=True) as a, \
                K.Simulator(mk(), dtype="c128", direct=False, support=False, defer=False) as b, \
                K.Simulator(mk(), dtype="c128", direct=False, defer=True) as c:
            pa = a.execute()
            pb = b.execute()
            pc = c.execute()
            sa, sb, sc = a.state(), b.state(), c.state()
        assert np.array_equal(pc, pb), (seed, alloc)
        assert np.array_equal(sc, sb), (seed, alloc)
        assert np.max(np.abs(pa - pb)) <= 1e-13, (seed, alloc)
        assert np.max(np.abs(sa - sb)) <= 1e-15, (seed, alloc)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_copy_ceiling_leaves_the_state_unchanged(gpu, dtype):
    """The copy-ceiling calibration copies every tile of every pass in place:
    one time per pass, and the state is bit for bit what it was."""
    say("copy ceiling", dtype)
    n = 18
    gates = W.layered(n, 4, 3)
    with K.Simulator(K.Circuit(n, gates), dtype=dtype) as sim:
        sim.execute()
        before = sim.state()
        ms = sim.copy_ceiling()
        after = sim.state()
        passes = sim.timings()["passes"]
    assert len(ms) == passes and np.all(ms > 0)
    assert np.array_equal(before, after)


SCATTER_GEN = {
    "layered20": (20, lambda n: W.layered(n, 3, 41)),
    "layered18": (18, lambda n: W.layered(n, 4, 5)),
    "mixed_tail": (16, lambda n: W.random_mixed(n, 100, 3) + [W.Gate(W.CNOT, (i + 1, i)) for i in range(n - 1)]),
    "haar_tail": (22, lambda n: W.haar_layers(n, 2, 3) + [W.Gate(W.CNOT, (i + 1, i)) for i in range(n - 1)]),
}
# (dtype, tile, low, circuit): shapes whose plans end with 2+ permutation passes
SCATTER_CASES = [("c128", 12, 3, "layered18"), ("c128", 12, 3, "mixed_tail"), ("c128", 10, 3, "layered20"),
                 ("c128", 10, 3, "mixed_tail"), ("c128", 11, 3, "haar_tail"), ("c64", 13, 4, "layered20"),
                 ("c64", 11, 3, "layered18"), ("c64", 11, 3, "mixed_tail")]


@pytest.mark.parametrize("dtype,tile,low,name", SCATTER_CASES, ids=["-".join(map(str, c)) for c in SCATTER_CASES])
def test_scatter_pass_bit_exact(gpu, dtype, tile, low, name):
    """The trailing permutation passes merged into one out-of-place pass
    (the state moves to the second buffer) give the same state bit for bit
    as the passes run in place, with the gate passes pinned (tile, low
    qubits, program order); and the oracle's to the usual tolerance."""
    n, gen = SCATTER_GEN[name]
    say("scatter", dtype, tile, low, name, n)
    gates = gen(n)
    opt = dict(tile_qubits=tile, low_qubits=low, reorder=False)
    _, _, text = K.Circuit(n, gates).plan(dtype, **opt)
    assert " scatter" in text, text.splitlines()[-1]
    a = state_of(K.Circuit(n, gates), dtype, **opt)
    b = state_of(K.Circuit(n, gates), dtype, scatter=False, **opt)
    assert np.array_equal(a, b)
    tol = 1e-12 if dtype == "c128" else 2e-6
    assert np.max(np.abs(a - O.simulate(n, gates))) <= tol


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_scatter_pass_every_read_path(gpu, dtype):
    """Around the buffer swap of the scatter pass: repeated execute(), the
    deferred form (the scatter pass run by the next read), proba_state,
    Z strings, the pass-by-pass run from a basis state and the copy-ceiling
    timing (which must leave the state as it was) all agree with the plan
    that runs its trailing passes in place."""
    say("scatter read paths", dtype)
    n = 19
    gates = W.layered(n, 3, 8)
    opt = dict(tile_qubits=12, low_qubits=3, reorder=False)
    masks = [1, 3, (1 << n) - 1, 1 << 7]
    with K.Simulator(K.Circuit(n, gates), dtype=dtype, **opt) as a, \
            K.Simulator(K.Circuit(n, gates), dtype=dtype, scatter=False, **opt) as b, \
            K.Simulator(K.Circuit(n, gates), dtype=dtype, defer=True, **opt) as d:
        pb = b.execute()
        sb = b.state()
        for _ in range(3):
            pa = a.execute()
        assert np.array_equal(pa, pb)
        assert np.array_equal(a.state(), sb)
        assert np.array_equal(a.expect_z(masks), b.expect_z(masks))
        assert np.array_equal(a.measure_proba(), b.measure_proba())
        pd = d.execute()
        assert np.max(np.abs(pd - pb)) <= (1e-12 if dtype == "c128" else 1e-5)
        assert np.array_equal(d.state(), sb)
        a.copy_ceiling()
        assert np.array_equal(a.state(), sb)
        a.init_basis(5)
        b.init_basis(5)
        a.run_timed()
        b.run_timed()
        assert np.array_equal(a.state(), b.state())


RELABEL_CASES = [
    ("layered", 20, "c128", 0, lambda n: W.layered(n, 8, 61)),
    ("layered", 20, "c64", 12, lambda n: W.layered(n, 8, 61)),
    ("brick", 18, "c128", 0, lambda n: W.brickwork(n, 10, 62, longrange_offset=9)),
    ("brick", 18, "c64", 12, lambda n: W.brickwork(n, 10, 62, longrange_offset=9)),
    ("mixed", 16, "c128", 10, lambda n: W.random_mixed(n, 250, 63, max_dense=4) + W.layered(n, 2, 64)),
    ("qft", 18, "c64", 12, lambda n: W.layered(n, 2, 5) + W.qft(n)),
    ("haar", 17, "c64", 11, lambda n: W.haar_layers(n, 12, 65)),
    ("haar", 18, "c128", 12, lambda n: W.haar_layers(n, 12, 65)),
    ("layered22", 22, "c128", 0, lambda n: W.layered(n, 10, 66)),
    ("layered22", 22, "c64", 13, lambda n: W.layered(n, 10, 66)),
]


@pytest.mark.parametrize("name,n,dtype,tile,gen", RELABEL_CASES,
                         ids=[f"{c[0]}-{c[2]}-{c[3]}" for c in RELABEL_CASES])
def test_relabelled_passes_bit_exact(gpu, name, n, dtype, tile, gen):
    """Relabelled passes (each reads its tile from the lowest state bits and
    writes out of place where the next pass reads) give the state of the
    in-place plan bit for bit, from |0...0> (execute) and pass by pass from a
    basis state; and the oracle's to the usual tolerance."""
    say("relabel", name, n, dtype, tile)
    gates = gen(n)
    _, _, text = K.Circuit(n, gates).plan(dtype, tile_qubits=tile)
    assert " relabel" in text
    # the same plan in place: the planner's low-qubit choice depends on the
    # relabelling, so both take the relabelled plan's
    opt = dict(tile_qubits=int(text.split("tile_bits=")[1].split()[0]), low_qubits=int(text.split("low=")[1].split()[0]))
    in_place = K.Circuit(n, gates).plan(dtype, relabel=False, **opt)[2]
    assert in_place.splitlines()[0] == K.Circuit(n, gates).plan(dtype, **opt)[2].splitlines()[0]
    with K.Simulator(K.Circuit(n, gates), dtype=dtype, **opt) as a, \
            K.Simulator(K.Circuit(n, gates), dtype=dtype, relabel=False, **opt) as b:
        pa, pb = a.execute(), b.execute()
        sa = a.state()
        assert np.array_equal(sa, b.state())
        assert np.array_equal(pa, pb)
        masks = [1, (1 << n) - 1, 0b1011]
        assert np.array_equal(a.expect_z(masks), b.expect_z(masks))
        a.init_basis(3)
        b.init_basis(3)
        a.run_timed()
        b.run_timed()
        assert np.array_equal(a.state(), b.state())
    tol = 1e-12 if dtype == "c128" else 2e-6
    assert np.max(np.abs(sa - O.simulate(n, gates))) <= tol
