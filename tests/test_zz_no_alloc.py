"""kate_state_expect_z_noalloc (the entry point that allocates nothing)
against kate_sim_expect_z on the same state, through the C ABI: two calls in
a row on one GPU, the second on a smaller state, and its argument checks."""
import ctypes as C

import numpy as np
import pytest

import kate_amd as K
from kate_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


class DeviceBuffer:
    """The caller's output buffer: hipMalloc / hipMemcpy of the HIP runtime
    libkate_hip.so runs on (the test allocates; the entry point does not)."""

    def __init__(self, nbytes):
        K.lib()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def read(self, count):
        out = np.zeros(count)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, count * 8, 2) == 0  # device to host
        return out

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.hip.hipFree(self.ptr)


def noalloc_z(sim, n, dtype, masks, d_out):
    m = np.array(masks, dtype=np.uint64)
    rc = K.lib().kate_state_expect_z_noalloc(0, K.KATE_C128 if dtype == "c128" else K.KATE_C64,
                                             C.c_void_p(sim.device_state()), n, m.ctypes.data_as(K.U64P), len(m),
                                             d_out.ptr, C.c_void_p(sim.stream()))
    assert rc == 0, K.lib().kate_last_error().decode()
    sim.synchronize()
    return d_out.read(len(m))


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_noalloc_expect_z_matches_simulator(gpu, dtype):
    with DeviceBuffer(16 * 8) as d_out:
        for n, seed in ((20, 1), (14, 2)):  # the second call on a smaller state
            say("noalloc", dtype, n)
            gates = W.random_mixed(n, 150, seed, max_dense=3) + W.layered(n, 2, seed)
            masks = [1 << q for q in range(min(n, 10))] + [int(x) for x in W.zstring_masks(n, 5, seed)]
            with K.Simulator(K.Circuit(n, gates), dtype=dtype) as sim:
                sim.execute()
                ref = sim.expect_z(masks)
                got = noalloc_z(sim, n, dtype, masks, d_out)
            say("max |noalloc - expect_z|", float(np.max(np.abs(got - ref))))
            assert np.max(np.abs(got - ref)) <= 1e-12


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    with DeviceBuffer(16 * 8) as d_out, K.Simulator(K.Circuit(14, W.layered(14, 1, 1)), dtype="c128") as sim:
        sim.execute()
        masks = np.arange(1, 17, dtype=np.uint64)
        rc = K.lib().kate_state_expect_z_noalloc(0, K.KATE_C128, C.c_void_p(sim.device_state()), 14,
                                                 masks.ctypes.data_as(K.U64P), 16, d_out.ptr,
                                                 C.c_void_p(sim.stream()))
        assert rc != 0
        assert "1..15 Z strings" in K.lib().kate_last_error().decode()
        rc = K.lib().kate_state_expect_z_noalloc(0, K.KATE_C128, C.c_void_p(sim.device_state()), 11,
                                                 masks.ctypes.data_as(K.U64P), 1, d_out.ptr,
                                                 C.c_void_p(sim.stream()))
        assert rc != 0 and "12.." in K.lib().kate_last_error().decode()
