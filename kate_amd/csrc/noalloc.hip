// noalloc.hip -- <Z_S> of a caller-owned device state without any device
// allocation (kate_state_expect_z_noalloc).
//
// kate_sim_expect_z reduces through the simulator's scratch buffer, which
// kate_sim_create allocates.  Callers that cannot allocate (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc) get
// the same reduction with its partial sums in a file-scope device array of
// fixed size: a grid of at most kBlocks workgroups strides over the state's
// 4096-amplitude chunks, each workgroup writes one row of kZ partial sums,
// and one workgroup adds the rows in a fixed order and divides by the norm.
// Every GPU holds its own copy of the array; its address is looked up once
// per device with hipGetSymbolAddress and only that address is passed to the
// kernels.  Calls that overlap on one GPU share the array and race.
#include <mutex>

#include "abi_util.hpp"
#include "kernels_dev.hpp"

namespace kate {
namespace {

constexpr int kBlocks = 2048;  // workgroups of the first level (rows of partial sums)
constexpr int kZ = 16;         // the norm and up to 15 Z strings

// rows of partial sums, kBlocks x kZ doubles (256 KiB)
__device__ __align__(256) double scratch[kBlocks * kZ];

struct ZMasks {
    uint64_t m[kZ];
};

// Level 1: workgroup b sums chunks b, b + gridDim.x, ... (4096 amplitudes
// each, amplitude it*256 + tid of a chunk per step) into kZ signed |a|^2
// sums, in a fixed order for a fixed grid.
template <typename R>
__global__ __launch_bounds__(256) void k_zgrid(const Cx<R>* __restrict__ st, uint64_t chunks, ZMasks mk, int m,
                                              double* __restrict__ rows)
{
    const uint32_t tid = threadIdx.x;
    double acc[kZ];
#pragma unroll
    for (int j = 0; j < kZ; j++) acc[j] = 0.0;
    for (uint64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        const uint64_t base = c << 12;
#pragma unroll 4
        for (int it = 0; it < 16; it++) {
            const uint64_t i = base | ((uint64_t)it << 8) | tid;
            const Cx<R> a = st[i];
            const double p = (double)a.re * (double)a.re + (double)a.im * (double)a.im;
#pragma unroll
            for (int j = 0; j < kZ; j++)
                if (j < m) acc[j] += (__popcll(i & mk.m[j]) & 1) ? -p : p;
        }
    }
    __shared__ double wpart[4][kZ];
#pragma unroll
    for (int j = 0; j < kZ; j++) acc[j] = wave_sum(acc[j]);
    if ((tid & 63u) == 0) {
#pragma unroll
        for (int j = 0; j < kZ; j++) wpart[tid >> 6][j] = acc[j];
    }
    __syncthreads();
    if (tid < (uint32_t)kZ) rows[(uint64_t)blockIdx.x * kZ + tid] = (wpart[0][tid] + wpart[1][tid]) + (wpart[2][tid] + wpart[3][tid]);
}

// Level 2 (one workgroup): column sums of `nrows` rows in a fixed order,
// then out[j - 1] = sum_j / sum_0 for j = 1..m-1.
__global__ __launch_bounds__(256) void k_zfinish(const double* __restrict__ rows, int nrows, int m,
                                                 double* __restrict__ out)
{
    __shared__ double lds[4];
    __shared__ double tot[kZ];
    for (int j = 0; j < m; j++) {
        double a = 0.0;
        for (int r = (int)threadIdx.x; r < nrows; r += 256) a += rows[(size_t)r * kZ + j];
        const double s = block_sum(a, lds);
        if (threadIdx.x == 0) tot[j] = s;
        __syncthreads();
    }
    if ((int)threadIdx.x >= 1 && (int)threadIdx.x < m) out[threadIdx.x - 1] = tot[threadIdx.x] / tot[0];
}

// Device address of `scratch` on each GPU (the host-side name is only a
// stand-in for the symbol): looked up once per device, with that device
// current.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

int scratch_on(int device, double** out)
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(KATE_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = static_cast<double*>(p);
    }
    *out = g_scratch[device];
    return KATE_OK;
}

}  // namespace
}  // namespace kate

extern "C" int kate_state_expect_z_noalloc(int32_t device, int32_t dtype, const void* d_state, int32_t nqubits,
                                           const uint64_t* masks, int32_t count, double* d_out, void* stream)
{
    using namespace kate;
    if (!d_state || !d_out || (count > 0 && !masks)) return set_error(KATE_ERR_ARG, "null argument");
    if (dtype != KATE_C128 && dtype != KATE_C64) return set_error(KATE_ERR_ARG, "dtype must be KATE_C128 or KATE_C64");
    if (nqubits < 12 || nqubits > KATE_MAX_QUBITS)
        return set_error(KATE_ERR_ARG, "kate_state_expect_z_noalloc takes states of 12.." +
                                           std::to_string(KATE_MAX_QUBITS) + " qubits");
    if (count < 1 || count > kZ - 1)
        return set_error(KATE_ERR_ARG, "kate_state_expect_z_noalloc takes 1.." + std::to_string(kZ - 1) +
                                           " Z strings per call (its scratch holds " + std::to_string(kZ) +
                                           " sums per row)");
    if (device < 0 || device >= kMaxDevices) return set_error(KATE_ERR_ARG, "device ordinal out of range");
    ZMasks mk{};
    for (int j = 0; j < count; j++) {
        if (nqubits < 64 && (masks[j] >> nqubits)) return set_error(KATE_ERR_ARG, "mask names a qubit outside the state");
        mk.m[j + 1] = masks[j];
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return set_error(KATE_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    double* rows = nullptr;
    if (int rc = scratch_on(device, &rows)) return rc;
    const uint64_t chunks = 1ull << (nqubits - 12);
    const unsigned grid = (unsigned)std::min<uint64_t>(chunks, (uint64_t)kBlocks);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KATE_C128)
        hipLaunchKernelGGL(k_zgrid<double>, dim3(grid), dim3(256), 0, s, (const Cx<double>*)d_state, chunks, mk,
                           count + 1, rows);
    else
        hipLaunchKernelGGL(k_zgrid<float>, dim3(grid), dim3(256), 0, s, (const Cx<float>*)d_state, chunks, mk,
                           count + 1, rows);
    hipLaunchKernelGGL(k_zfinish, dim3(1), dim3(256), 0, s, (const double*)rows, (int)grid, count + 1, d_out);
    e = hipGetLastError();
    if (e != hipSuccess) return set_error(KATE_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
    return KATE_OK;
}
