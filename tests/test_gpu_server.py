"""The continuous-batching server: results against the oracle, each result
bitwise the same alone and in a mix, admission bounded by the HBM budget."""
import numpy as np
import pytest

import kate_amd as K
from kate_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_requests(count, seed, nmin=12, nmax=20):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        masks = [int(m) for m in W.zstring_masks(n, 8, seed * 100 + i)]
        if i % 2 == 0:
            gates = W.layered(n, int(rng.integers(2, 6)), seed + i)
            tp = None
        else:
            gates = W.qft(n)
            tp = W.basis_input(n, int(rng.integers(0, 1 << n)))
        reqs.append((n, gates, tp, masks))
    return reqs


def oracle_z(n, gates, tp, masks):
    st = O.simulate(n, gates, tp)
    return O.expect_z(n, st, masks) / O.expect_z(n, st, [0])[0]


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 2e-6)])
def test_server_results_match_oracle(gpu, dtype, tol):
    say("server", dtype)
    reqs = make_requests(10, 3)
    with K.Server(dtype=dtype, streams=4, max_active=8, hbm_budget_bytes=2 << 30) as srv:
        ids = [srv.submit(K.Circuit(n, g), m, tp) for n, g, tp, m in reqs]
        outs = [srv.wait(i) for i in ids]
    for (n, g, tp, m), (z, st) in zip(reqs, outs):
        assert st["nqubits"] == n and st["latency_ms"] > 0
        assert np.max(np.abs(z - oracle_z(n, g, tp, m))) < tol


def test_server_alone_equals_mixed(gpu):
    say("server alone/mixed")
    reqs = make_requests(12, 7, 12, 22)
    probe = reqs[5]
    with K.Server(dtype="c64", streams=4, max_active=16, hbm_budget_bytes=2 << 30) as srv:
        alone, _ = srv.wait(srv.submit(K.Circuit(probe[0], probe[1]), probe[3], probe[2]))
    with K.Server(dtype="c64", streams=4, max_active=16, hbm_budget_bytes=2 << 30) as srv:
        ids = [srv.submit(K.Circuit(n, g), m, tp) for n, g, tp, m in reqs]
        mixed = [srv.wait(i)[0] for i in ids]
    assert np.array_equal(alone, mixed[5])


def test_server_admission_by_footprint(gpu):
    say("server admission")
    reqs = [(20, W.layered(20, 3, i), None, [1, 2, 3]) for i in range(8)]  # 16 MiB FP64 states
    with K.Server(dtype="c128", streams=4, max_active=32, hbm_budget_bytes=40 << 20) as srv:
        ids = [srv.submit(K.Circuit(n, g), m, tp) for n, g, tp, m in reqs]
        seen_queue = srv.pending()["queued"] > 0
        outs = [srv.wait(i) for i in ids]
        assert srv.pending()["completed"] == len(reqs)
    assert seen_queue  # only two 16 MiB states fit a 40 MiB arena at a time
    for (n, g, tp, m), (z, _) in zip(reqs, outs):
        assert np.max(np.abs(z - oracle_z(n, g, tp, m))) < 1e-12


def test_server_rejects_request_larger_than_budget(gpu):
    say("server oversize")
    with K.Server(dtype="c128", hbm_budget_bytes=8 << 20) as srv:
        rid = srv.submit(K.Circuit(22, W.layered(22, 1, 0)), [1])
        with pytest.raises(K.KateError, match="does not fit"):
            srv.wait(rid)


def test_server_28_qubit_request_alone_equals_mixed(gpu):
    """A 28-qubit request (2 GiB complex64 state) served alone and amid
    smaller requests gives the same <Z_S> bit for bit, and matches the same
    circuit in complex128 on a simulator to FP32 accuracy."""
    say("server 28q alone/mixed")
    n = 28
    gates = W.layered(n, 8, 77)
    masks = [int(m) for m in W.zstring_masks(n, 8, 78)]
    small = make_requests(10, 9, 12, 22)
    budget = 6 << 30
    with K.Server(dtype="c64", streams=4, max_active=16, hbm_budget_bytes=budget) as srv:
        alone, st = srv.wait(srv.submit(K.Circuit(n, gates), masks))
    assert st["nqubits"] == n
    with K.Server(dtype="c64", streams=4, max_active=16, hbm_budget_bytes=budget) as srv:
        ids = [srv.submit(K.Circuit(m_, g), mk, tp) for m_, g, tp, mk in small[:5]]
        big = srv.submit(K.Circuit(n, gates), masks)
        ids += [srv.submit(K.Circuit(m_, g), mk, tp) for m_, g, tp, mk in small[5:]]
        mixed, _ = srv.wait(big)
        rest = [srv.wait(i)[0] for i in ids]
    assert np.array_equal(alone, mixed)
    for (m_, g, tp, mk), z in zip(small, rest):
        assert np.max(np.abs(z - oracle_z(m_, g, tp, mk))) < 2e-6
    with K.Simulator(K.Circuit(n, gates), dtype="c128") as ref:
        ref.execute()
        z64 = ref.expect_z(masks)
    say("28q FP32 server vs FP64 max |dZ|", float(np.max(np.abs(alone - z64))))
    assert np.max(np.abs(alone - z64)) < 2e-6


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 2e-6)])
def test_server_out_of_place_requests(gpu, dtype, tol):
    """Requests of 22 qubits or more run on two state buffers of their own
    with relabelled passes (each reads its tile from the lowest state bits
    and writes where the next pass reads): layered circuits from |0...0> and
    QFTs on basis inputs, served beside small in-place requests, match the
    oracle."""
    say("server out-of-place", dtype)
    n = 22
    big = [(n, W.layered(n, 8, 93), None, [int(m) for m in W.zstring_masks(n, 8, 94)]),
           (n, W.qft(n), W.basis_input(n, 0x2D5A3), [int(m) for m in W.zstring_masks(n, 8, 95)]),
           (n + 1, W.layered(n + 1, 6, 96), None, [int(m) for m in W.zstring_masks(n + 1, 8, 97)]),
           (n, W.qft(n), W.basis_input(n, 0x1F0F1), [int(m) for m in W.zstring_masks(n, 8, 98)])]
    for m_, g, tp, _ in big:
        if tp is None:  # from |0...0>: relabelled after the support region
            assert " relabel" in K.Circuit(m_, g).plan(dtype, tile_qubits=12)[2]
    small = make_requests(4, 92, 12, 16)
    reqs = [r for pair in zip(big, small) for r in pair]
    with K.Server(dtype=dtype, streams=4, max_active=8) as srv:
        ids = [srv.submit(K.Circuit(n, g), m, tp) for n, g, tp, m in reqs]
        outs = [srv.wait(i) for i in ids]
    for (n, g, tp, m), (z, st) in zip(reqs, outs):
        if n >= 22:  # a second state buffer in the request's footprint
            assert st["footprint_bytes"] >= 2 * (1 << n) * (16 if dtype == "c128" else 8)
        assert np.max(np.abs(z - oracle_z(n, g, tp, m))) < tol
