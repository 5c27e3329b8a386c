"""Kronecker steps in Euler form (each composed 1-qubit gate written as
D1 * RX * D2, the diagonals of a step applied as one table) against the
oracle and against the same plan with dense 2x2 products: RX/RZ layers (no
right diagonal) and Haar-random 1-qubit layers (with it), complex128 at both
register widths and complex64."""
import numpy as np
import pytest

import kate_amd as K
from kate_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = {"c128": 1e-12, "c64": 2e-6}


def say(*a):
    print(*a, flush=True)


def state_of(c, dtype, **opt):
    with K.Simulator(c, dtype=dtype, **opt) as sim:
        sim.execute()
        return sim.state()


CIRCUITS = {
    "layered": lambda n: W.layered(n, 4, 11),
    "haar": lambda n: W.haar_layers(n, 4, 12),
    "mixed": lambda n: W.haar_layers(n, 2, 13) + W.random_mixed(n, 80, 14, max_dense=3) + W.layered(n, 2, 15),
}


@pytest.mark.parametrize("dtype,mb", [("c128", 3), ("c128", 4), ("c64", 4)])
@pytest.mark.parametrize("name", sorted(CIRCUITS))
@pytest.mark.parametrize("n", [12, 14])
def test_euler_steps_match_oracle(gpu, monkeypatch, dtype, mb, name, n):
    say("euler", dtype, mb, name, n)
    gates = CIRCUITS[name](n)
    monkeypatch.setenv("KATE_DESCRIBE_OPS", "1")
    _, _, text = K.Circuit(n, gates).plan(dtype, micro_bits=mb)
    monkeypatch.delenv("KATE_DESCRIBE_OPS")
    assert " ekron:" in text  # the case exercises the Euler-form op
    ref = O.simulate(n, gates)
    st = state_of(K.Circuit(n, gates), dtype, micro_bits=mb)
    assert np.max(np.abs(st - ref)) <= TOL[dtype]
    plain = state_of(K.Circuit(n, gates), dtype, micro_bits=mb, euler=False)
    assert np.max(np.abs(st - plain)) <= (1e-14 if dtype == "c128" else 1e-6)


def test_euler_at_20_qubits(gpu):
    say("euler 20q")
    n = 20
    for gates in (W.layered(n, 8, 16), W.haar_layers(n, 6, 17)):
        a = state_of(K.Circuit(n, gates), "c128")
        b = state_of(K.Circuit(n, gates), "c128", euler=False)
        assert np.max(np.abs(a - b)) <= 1e-14
        assert abs(np.linalg.norm(a) - 1.0) <= 1e-12
