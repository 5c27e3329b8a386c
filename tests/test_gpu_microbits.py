"""complex128 with both register micro-step widths (8 or 16 amplitudes per
thread) against the oracle and Kate's fixtures: random circuits with every
gate kind (4-qubit dense gates leave the 3-bit register path), every tile
shape, Kate's config-1 circuit, swap programs, and batches."""
import numpy as np
import pytest

import kate_amd as K
from fixture_io import unpack_gates
from kate_amd import workloads as W
from oracle import oracle as O
from util import compiled_from_circuit, fixtures, load, to_real_order

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def say(*a):
    print(*a, flush=True)


def state_of(c, **opt):
    with K.Simulator(c, dtype="c128", **opt) as sim:
        pr = sim.execute()
        return sim.state(real_order=False), pr


@pytest.mark.parametrize("mb", [3, 4])
@pytest.mark.parametrize("n,ngates,maxd,seed,tile", [(5, 40, 3, 1, 10), (12, 200, 4, 4, 12), (14, 200, 4, 5, 11),
                                                     (16, 150, 3, 6, 13), (17, 160, 4, 8, 12), (13, 60, 6, 7, 10)])
def test_random_circuits(gpu, mb, n, ngates, maxd, seed, tile):
    say("microbits random", mb, n, ngates, maxd, seed, tile)
    gates = W.random_mixed(n, ngates, seed, max_dense=maxd) + W.layered(n, 2, seed)
    st, _ = state_of(K.Circuit(n, gates), tile_qubits=tile, micro_bits=mb)
    assert np.max(np.abs(st - O.simulate(n, gates))) <= TOL64


@pytest.mark.parametrize("mb", [3, 4])
def test_config1_against_kate(gpu, mb):
    say("microbits config1", mb)
    z, _ = load(fixtures("kate_config1_layered20.npz")[0])
    gates = unpack_gates(z, "in_")
    st, pr = state_of(K.Circuit(20, gates), micro_bits=mb)
    assert np.max(np.abs(st[z["sample_index"]] - z["sample_amp"])) <= TOL64
    assert np.max(np.abs(pr - z["proba"])) < 1e-9


@pytest.mark.parametrize("mb", [3, 4])
def test_swap_programs(gpu, mb):
    say("microbits swaps", mb)
    n = 15
    gates = W.random_mixed(n, 150, 63)
    c = K.Circuit(n, gates).schedule().group(6).allocate(3)
    prog = compiled_from_circuit(c)
    with K.Simulator(c, dtype="c128", micro_bits=mb) as sim:
        sim.execute()
        st = sim.state(real_order=True)
    assert np.max(np.abs(st - to_real_order(O.run(prog), prog.final_inverse_permutation))) <= TOL64


@pytest.mark.parametrize("mb", [3, 4])
def test_batch(gpu, mb):
    say("microbits batch", mb)
    work = W.batch_workload(12, seed=21, nmin=10, nmax=16, dmin=4, dmax=10)
    work.append((15, W.random_mixed(15, 80, 5, max_dense=4)))
    with K.Batch([K.Circuit(n, g) for n, g in work], dtype="c128", micro_bits=mb) as b:
        b.run()
        z, _ = b.expect_z()
    for (n, gates), zc in zip(work, z):
        ref = O.expect_z(n, O.simulate(n, gates), [1 << q for q in range(n)])
        assert np.max(np.abs(zc - ref)) < 1e-12


def test_micro_bits_default_is_valid(gpu):
    n = 13
    gates = W.random_mixed(n, 60, 9)
    a, _ = state_of(K.Circuit(n, gates), micro_bits=3)
    b, _ = state_of(K.Circuit(n, gates), micro_bits=4)
    assert np.max(np.abs(a - b)) <= TOL64
