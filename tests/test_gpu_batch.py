"""Batched runs: many circuits in one persistent launch, <Z_q> per circuit,
against the FP64 oracle; graph replays against eager runs."""
import numpy as np
import pytest

import kate_amd as K
from kate_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def reference_z(n, gates):
    st = O.simulate(n, gates)
    return O.expect_z(n, st, [1 << q for q in range(n)])


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 2e-6)])
@pytest.mark.parametrize("tile", [12, 13])
def test_batch_matches_oracle(gpu, dtype, tol, tile):
    say("batch", dtype, tile)
    work = W.batch_workload(24, seed=5, nmin=10, nmax=17, dmin=4, dmax=12)
    work.append((16, W.layered(16, 3, 2)))  # multi-pass with layout CNOTs
    circuits = [K.Circuit(n, g) for n, g in work]
    with K.Batch(circuits, dtype=dtype, tile_qubits=tile) as b:
        info = b.info()
        assert info["circuits"] == len(work)
        b.run()
        z, norms = b.expect_z()
    for (n, gates), zc, nm in zip(work, z, norms):
        ref = reference_z(n, gates)
        assert np.max(np.abs(zc - ref)) < tol, (n, np.max(np.abs(zc - ref)))
        assert abs(nm - 1) < (1e-12 if dtype == "c128" else 1e-5)


def test_batch_fused_programs(gpu):
    say("batch fused")
    work = W.batch_workload(16, seed=9, nmin=10, nmax=16, dmin=6, dmax=12)
    circuits = [K.Circuit(n, g).schedule().fuse(4) for n, g in work]
    with K.Batch(circuits, dtype="c64") as b:
        b.run()
        z, _ = b.expect_z()
    for (n, gates), zc in zip(work, z):
        assert np.max(np.abs(zc - reference_z(n, gates))) < 2e-6


def test_batch_graph_replays_equal_eager(gpu):
    say("batch graph")
    work = W.batch_workload(32, seed=3, nmin=10, nmax=16, dmin=6, dmax=16)
    circuits = [K.Circuit(n, g) for n, g in work]
    with K.Batch(circuits, dtype="c64") as b:
        b.run()
        eager, en = b.expect_z()
        b.capture()
        for _ in range(3):
            b.run()
            z, nr = b.expect_z()
            assert all(np.array_equal(x, y) for x, y in zip(z, eager))
            assert np.array_equal(nr, en)
            # eager work from the program in between replays
            with K.Simulator(K.Circuit(12, W.random_mixed(12, 30, 1))) as sim:
                sim.execute()


def test_batch_result_independent_of_company(gpu):
    """A circuit's <Z> is bitwise the same alone and inside a bigger batch."""
    say("batch company")
    work = W.batch_workload(20, seed=13, nmin=10, nmax=16, dmin=6, dmax=14)
    alone = []
    for n, g in work[:3]:
        with K.Batch([K.Circuit(n, g)], dtype="c64") as b:
            b.run()
            alone.append(b.expect_z()[0][0])
    with K.Batch([K.Circuit(n, g) for n, g in work], dtype="c64") as b:
        b.run()
        z, _ = b.expect_z()
    for a, m in zip(alone, z[:3]):
        assert np.array_equal(a, m)


def test_batch_compiled_circuits_report_real_qubits(gpu):
    """Circuits compiled by Kate's passes (scheduling, grouping and the
    allocate() swap plan, which relabels qubits) give <Z_q> per real qubit q,
    as the oracle computes it on the uncompiled circuit."""
    say("batch compiled")
    work = W.batch_workload(12, seed=21, nmin=12, nmax=16, dmin=6, dmax=12)
    work.append((14, W.layered(14, 3, 5)))
    circuits = [K.Circuit(n, g).schedule().group(8).allocate(2) for n, g in work]
    assert any(c.permutations()[1] != list(range(c.n)) for c in circuits)
    for dtype, tol in (("c128", 1e-12), ("c64", 2e-6)):
        with K.Batch(circuits, dtype=dtype) as b:
            b.run()
            z, _ = b.expect_z()
        for (n, gates), zc in zip(work, z):
            err = np.max(np.abs(zc - reference_z(n, gates)))
            assert err < tol, (dtype, n, err)


def test_full_c4_batch_with_autotuned_shape(gpu):
    """The full batched workload (4096 Haar brickwork circuits, n
    log-uniform in [10, 20], depth 8..32, complex64) in one persistent
    launch, in the shape the autotuner picks on a sample: <Z_q> of 64 seeded
    circuits against the FP64 oracle, and every norm."""
    from kate_amd.autotune import autotune_batch, build

    say("full C4: autotune")
    work = W.batch_workload(4096, seed=4)
    best, table = autotune_batch(work, dtype="c64", sample=64, reps=1)
    say("full C4: picked", best)
    with K.Batch(build(work, best["k"]), dtype="c64", tile_qubits=best["tile"], low_qubits=best["low"],
                 wgs_per_cu=best["wgs_per_cu"]) as b:
        b.capture()
        b.run()
        z, norms = b.expect_z()
    assert np.max(np.abs(norms - 1.0)) < 1e-5
    rng = np.random.default_rng(7)
    idx = rng.choice(len(work), size=64, replace=False)
    worst = 0.0
    for j, i in enumerate(idx):
        n, gates = work[int(i)]
        err = float(np.max(np.abs(z[int(i)] - reference_z(n, gates))))
        worst = max(worst, err)
        if j % 16 == 15:
            say(f"full C4: {j + 1} of 64 circuits checked, worst <Z> error {worst:.2e}")
        assert err < 2e-6, (int(i), n, err)
