// kernels.hip -- initialisation, reduction and batch-driver kernels and the
// launchers of every kernel (the pass and batch kernels themselves are
// instantiated in pass_f64.hip, pass_f32.hip, batch_f64.hip, batch_f32.hip).
#include "kernels_dev.hpp"
#include "kernels_inst.hpp"

namespace kate {
namespace {


// Fills n words with v (a kernel, so that it can sit in a captured graph).
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// |0...0> for every circuit of a batch.
template <typename R>
__global__ __launch_bounds__(256) void k_batch_init(Cx<R>* __restrict__ states, uint64_t total,
                                                    const uint64_t* __restrict__ circ_off, uint32_t count)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride)
        states[i] = Cx<R>{(R)0, (R)0};
}

template <typename R>
__global__ __launch_bounds__(256) void k_batch_ones(Cx<R>* __restrict__ states, const uint64_t* __restrict__ circ_off,
                                                    uint32_t count)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < count) states[circ_off[c]] = Cx<R>{(R)1, (R)0};
}

// Per circuit (one workgroup each): norm and <Z_q> for every qubit, FP64,
// fixed summation order.  Amplitude i = it*256 + tid.
template <typename R>
__global__ __launch_bounds__(256) void k_batch_z(const Cx<R>* __restrict__ states, const uint64_t* __restrict__ circ_off,
                                                 const int32_t* __restrict__ nq, const int32_t* __restrict__ nq_int,
                                                 const int64_t* __restrict__ zoff, double* __restrict__ z,
                                                 double* __restrict__ norms)
{
    __shared__ double lds[4];
    const uint32_t c = blockIdx.x, tid = threadIdx.x;
    const Cx<R>* st = states + circ_off[c];
    const int n = nq[c];
    const uint64_t iters = 1ull << (nq_int[c] - 8);
    double tot = 0.0, hi[24];
#pragma unroll
    for (int q = 0; q < 24; q++) hi[q] = 0.0;
    for (uint64_t it = 0; it < iters; it++) {
        const Cx<R> a = st[(it << 8) | tid];
        const double p = (double)a.re * (double)a.re + (double)a.im * (double)a.im;
        tot += p;
#pragma unroll
        for (int q = 0; q < 24; q++)
            if (q + 8 < n && ((it >> q) & 1ull)) hi[q] += p;
    }
    const double total = block_sum(tot, lds);
    double* out = z + zoff[c];
    for (int q = 0; q < n; q++) {
        double s1;
        if (q < 8) {
            s1 = block_sum(((tid >> q) & 1u) ? tot : 0.0, lds);
        } else {
            double h = 0.0;
#pragma unroll
            for (int j = 0; j < 24; j++)
                if (j == q - 8) h = hi[j];
            s1 = block_sum(h, lds);
        }
        if (tid == 0) out[q] = (total - 2.0 * s1) / total;
    }
    if (tid == 0) norms[c] = total;
}

// ---- initialisation --------------------------------------------------------

template <typename R>
__global__ __launch_bounds__(256) void k_init_basis(Cx<R>* __restrict__ st, uint64_t n, uint64_t index)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        st[i] = Cx<R>{(R)(i == index ? 1 : 0), (R)0};
}

// Product state, multiplied qubit by qubit in Kate's order
// (initialize_probastate, HIPcode/simulator.hpp:26-38).
template <typename R>
__global__ __launch_bounds__(256) void k_init_product(Cx<R>* __restrict__ st, uint64_t n, int nq,
                                                       const double* __restrict__ f)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double re = 1.0, im = 0.0;
        for (int j = 0; j < nq; j++) {
            const double* c = f + 2 * (((i >> j) & 1ull) * (uint64_t)nq + (uint64_t)j);
            const double t = re * c[0] - im * c[1];
            im = re * c[1] + im * c[0];
            re = t;
        }
        st[i] = Cx<R>{(R)re, (R)im};
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_scale(Cx<R>* __restrict__ st, uint64_t n, double f)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        Cx<R> a = st[i];
        st[i] = Cx<R>{(R)((double)a.re * f), (R)((double)a.im * f)};
    }
}

// ---- deterministic reductions ------------------------------------------------

// Level 0 of the amplitude sums: a block reduces 4096 amplitudes (tile index
// it*256 + tid) to one record {total, S1[bit 0..11]}.
// With GATHER the amplitude of index i is read at offset M(i), M linear over
// GF(2) with columns gm.col[] (a state whose last permutation passes have
// not run yet: M maps each final index to where its amplitude still is).
struct GatherMap {
    uint64_t col[kMaxQubits];
};

template <typename R, bool GATHER>
__global__ __launch_bounds__(256) void k_sums0(const Cx<R>* __restrict__ st, double* __restrict__ rec, GatherMap gm)
{
    const uint32_t tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x << 12;
    uint64_t gbase = 0;
    if constexpr (GATHER) {
        const uint32_t blk = blockIdx.x;
#pragma unroll
        for (int j = 0; j < kMaxQubits - 12; j++)
            if ((blk >> j) & 1u) gbase ^= gm.col[12 + j];
#pragma unroll
        for (int j = 0; j < 8; j++)
            if ((tid >> j) & 1u) gbase ^= gm.col[j];
    }
    double tr = 0, ti = 0, hr[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0};
#pragma unroll
    for (int it = 0; it < 16; it++) {
        uint64_t off = base + ((uint64_t)it << 8) + tid;
        if constexpr (GATHER) {
            off = gbase;
#pragma unroll
            for (int j = 0; j < 4; j++)
                if ((it >> j) & 1) off ^= gm.col[8 + j];
        }
        const Cx<R> a = st[off];
        tr += (double)a.re;
        ti += (double)a.im;
#pragma unroll
        for (int b = 0; b < 4; b++)
            if ((it >> b) & 1) {
                hr[b] += (double)a.re;
                hi[b] += (double)a.im;
            }
    }
    // Per wave (butterflies, no barrier): the totals, the sums over lanes
    // with lane bit q set (q < 6) and the loop-bit sums; one barrier, then
    // thread 0 adds the four waves in a fixed order.  Bits 6 and 7 of the
    // amplitude index are the wave index, so their sums are the totals of
    // waves 1, 3 and of waves 2, 3.
    constexpr int V = 2 + 12 + 8;
    double val[V];
    val[0] = wave_sum(tr);
    val[1] = wave_sum(ti);
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const bool on = (tid >> q) & 1u;
        val[2 + 2 * q] = wave_sum(on ? tr : 0.0);
        val[3 + 2 * q] = wave_sum(on ? ti : 0.0);
    }
#pragma unroll
    for (int b = 0; b < 4; b++) {
        val[14 + 2 * b] = wave_sum(hr[b]);
        val[15 + 2 * b] = wave_sum(hi[b]);
    }
    __shared__ double part[4][V];
    const uint32_t w = tid >> 6;
    if ((tid & 63u) == 0) {
#pragma unroll
        for (int j = 0; j < V; j++) part[w][j] = val[j];
    }
    __syncthreads();
    if (tid == 0) {
        double* out = rec + (uint64_t)blockIdx.x * 26;
        auto all = [&](int j) { return (part[0][j] + part[1][j]) + (part[2][j] + part[3][j]); };
        out[0] = all(0);
        out[1] = all(1);
        for (int q = 0; q < 6; q++) {
            out[2 + 2 * q] = all(2 + 2 * q);
            out[3 + 2 * q] = all(3 + 2 * q);
        }
        out[14] = part[1][0] + part[3][0];  // index bit 6
        out[15] = part[1][1] + part[3][1];
        out[16] = part[2][0] + part[3][0];  // index bit 7
        out[17] = part[2][1] + part[3][1];
        for (int b = 0; b < 4; b++) {
            out[18 + 2 * b] = all(14 + 2 * b);
            out[19 + 2 * b] = all(15 + 2 * b);
        }
    }
}

// Level >= 1: a block reduces 2^C records {total, S1[0..K)} (one per thread)
// into {total, S1[0..K), S1[K..K+C)}.
__global__ __launch_bounds__(256) void k_sums_rec(const double* __restrict__ in, int K, int C, double* __restrict__ out)
{
    __shared__ double lds[4];
    const uint32_t tid = threadIdx.x;
    const bool valid = tid < (1u << C);
    const uint64_t r = ((uint64_t)blockIdx.x << C) + tid;
    const int win = 2 * (1 + K), wout = 2 * (1 + K + C);
    const double* src = in + r * (uint64_t)win;
    double* dst = out + (uint64_t)blockIdx.x * (uint64_t)wout;
    const double tr = valid ? src[0] : 0.0, ti = valid ? src[1] : 0.0;
    for (int f = 0; f < win; f++) {
        const double s = block_sum(valid ? src[f] : 0.0, lds);
        if (tid == 0) dst[f] = s;
    }
    for (int c = 0; c < C; c++) {
        const bool on = valid && ((tid >> c) & 1u);
        double s = block_sum(on ? tr : 0.0, lds);
        if (tid == 0) dst[win + 2 * c] = s;
        s = block_sum(on ? ti : 0.0, lds);
        if (tid == 0) dst[win + 2 * c + 1] = s;
    }
}

// Level 0 of the Z sums: a block reduces 4096 amplitudes to MB signed sums.
template <typename R, int MB>
__global__ __launch_bounds__(256) void k_z0(const Cx<R>* __restrict__ st, const uint64_t* __restrict__ masks, int m,
                                            double* __restrict__ part)
{
    const uint32_t tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x << 12;
    double acc[MB];
    uint32_t sgn[MB];
    uint64_t mk[MB];
#pragma unroll
    for (int j = 0; j < MB; j++) {
        acc[j] = 0.0;
        mk[j] = j < m ? masks[j] : 0ull;
        sgn[j] = (uint32_t)(__popcll((base | tid) & mk[j]) & 1);
    }
#pragma unroll
    for (int it = 0; it < 16; it++) {
        const Cx<R> a = st[base + ((uint64_t)it << 8) + tid];
        const double p = (double)a.re * (double)a.re + (double)a.im * (double)a.im;
#pragma unroll
        for (int j = 0; j < MB; j++) {
            const uint32_t f = sgn[j] ^ (uint32_t)(__popcll(((uint64_t)it << 8) & mk[j]) & 1);
            acc[j] += f ? -p : p;
        }
    }
    // wave butterflies, then one barrier and a fixed-order sum of the waves
    __shared__ double wpart[4][MB];
#pragma unroll
    for (int j = 0; j < MB; j++) acc[j] = wave_sum(acc[j]);
    if ((tid & 63u) == 0) {
#pragma unroll
        for (int j = 0; j < MB; j++) wpart[tid >> 6][j] = acc[j];
    }
    __syncthreads();
    if (tid < (uint32_t)MB) {
        double* out = part + (uint64_t)blockIdx.x * MB;
        out[tid] = (wpart[0][tid] + wpart[1][tid]) + (wpart[2][tid] + wpart[3][tid]);
    }
}

// Level 0 of <a|b> = sum conj(a_i) b_i (FP64): a block reduces 4096
// amplitude pairs to {re, im, |a|^2, |b|^2}.
// Also folds max |a_i - b_i| into *maxabs (bits of a non-negative double,
// which order like the values; the max does not depend on the order).
template <typename RA, typename RB>
__global__ __launch_bounds__(256) void k_overlap0(const Cx<RA>* __restrict__ a, const Cx<RB>* __restrict__ b,
                                                  double* __restrict__ part, unsigned long long* __restrict__ maxabs)
{
    __shared__ double lds[4];
    const uint64_t base = (uint64_t)blockIdx.x << 12;
    double re = 0, im = 0, na = 0, nb = 0, mx = 0;
#pragma unroll
    for (int it = 0; it < 16; it++) {
        const uint64_t i = base + ((uint64_t)it << 8) + threadIdx.x;
        const Cx<RA> x = a[i];
        const Cx<RB> y = b[i];
        const double xr = x.re, xi = x.im, yr = y.re, yi = y.im;
        re += xr * yr + xi * yi;
        im += xr * yi - xi * yr;
        na += xr * xr + xi * xi;
        nb += yr * yr + yi * yi;
        const double dr = xr - yr, di = xi - yi;
        mx = fmax(mx, sqrt(dr * dr + di * di));
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(maxabs, (unsigned long long)__double_as_longlong(mx));
    double* out = part + (uint64_t)blockIdx.x * 4;
    double v = block_sum(re, lds);
    if (threadIdx.x == 0) out[0] = v;
    v = block_sum(im, lds);
    if (threadIdx.x == 0) out[1] = v;
    v = block_sum(na, lds);
    if (threadIdx.x == 0) out[2] = v;
    v = block_sum(nb, lds);
    if (threadIdx.x == 0) out[3] = v;
}

// rows[count][W] -> out[ceil(count/4096)][W] by fixed-order block sums.
__global__ __launch_bounds__(256) void k_rowsum(const double* __restrict__ rows, uint64_t count, int W,
                                                double* __restrict__ out)
{
    __shared__ double lds[4];
    const uint64_t r0 = (uint64_t)blockIdx.x << 12;
    for (int w = 0; w < W; w++) {
        double a = 0.0;
        for (int i = 0; i < 16; i++) {
            const uint64_t r = r0 + ((uint64_t)i << 8) + threadIdx.x;
            if (r < count) a += rows[r * W + w];
        }
        const double s = block_sum(a, lds);
        if (threadIdx.x == 0) out[(uint64_t)blockIdx.x * W + w] = s;
    }
}

// Calibration: the pass's tiles read and written back in place with no
// arithmetic (every amplitude once each way, the load positions of the
// pass, 2^(TB-4) threads of 16 registers, no LDS).  `one` is 1 at run time,
// so the stores cannot be elided.
template <typename R, int TB>
__global__ __launch_bounds__(1 << (TB - 4)) void k_copy_tile(Cx<R>* __restrict__ state, PassArgs pa, R one)
{
    constexpr int NB = 4;
    const uint32_t tid = threadIdx.x;
    uint64_t tb = rest_base(pa, blockIdx.x);
#pragma unroll
    for (int b = 0; b < TB - NB; b++) tb |= (uint64_t)((tid >> b) & 1u) << pa.load_pos[b];
    uint64_t off[1 << NB];
    Cx<R> v[1 << NB];
#pragma unroll
    for (int i = 0; i < (1 << NB); i++) {
        uint64_t o = tb;
#pragma unroll
        for (int k = 0; k < NB; k++)
            if ((i >> k) & 1) o |= 1ull << pa.load_pos[TB - NB + k];
        off[i] = o;
        v[i] = state[o];
    }
#pragma unroll
    for (int i = 0; i < (1 << NB); i++) state[off[i]] = Cx<R>{v[i].re * one, v[i].im * one};
}

// Out-of-place permutation pass (Plan::scatter_pass): tile `blk` of `src`
// (load_pos / rest_pos, read in runs as in k_pass, made from a basis state
// or read through the zero qubits in a run from |0...0>) goes to LDS; then
// store index u (thread bits, then register bits) reads LDS at the XOR of
// store_byte[k] over its set bits and writes dst at the XOR of out_tile[k]
// and of out_rest[i] over the set bits of blk.  Store index bits 0..low-1
// are state offset bits 0..low-1, so each wave writes whole runs.
template <typename R, int TB>
__global__ __launch_bounds__(1 << (TB - 4)) void k_scatter(const Cx<R>* __restrict__ src, Cx<R>* __restrict__ dst,
                                                          PassArgs pa, const ScatterArgs* __restrict__ sa)
{
    constexpr int NB = 4, N = 1 << NB, NT = 1 << (TB - NB);
    __shared__ Cx<R> tile[1 << TB];
    const uint32_t tid = threadIdx.x;
    const uint64_t blk = blockIdx.x;
    uint64_t tb = rest_base(pa, blk);
#pragma unroll
    for (int b = 0; b < TB - NB; b++) tb |= (uint64_t)((tid >> b) & 1u) << pa.load_pos[b];
    Cx<R> r[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        uint64_t off = tb;
#pragma unroll
        for (int k = 0; k < NB; k++)
            if ((i >> k) & 1) off |= 1ull << pa.load_pos[TB - NB + k];
        r[i] = pa.synth_basis ? synth_amp<R>(pa, off) : ((off & pa.zero_mask) ? Cx<R>{(R)0, (R)0} : src[off]);
    }
    const uint32_t tsw = swz<R>(tid) * (uint32_t)sizeof(Cx<R>);
#pragma unroll
    for (int i = 0; i < N; i++) lds_at(tile, tsw ^ (swz<R>((uint32_t)i * NT) * (uint32_t)sizeof(Cx<R>))) = r[i];
    __syncthreads();
    scatter_store<R, TB, NB>(dst, pa, blk, tile, sa);
}

template <typename R>
int scatter_r(int tile_bits, const void* src, void* dst, const PassArgs& pa, const ScatterArgs* sa, hipStream_t s)
{
    const unsigned grid = 1u << pa.nrest;
    switch (tile_bits) {
#define SC(TB) case TB: hipLaunchKernelGGL((k_scatter<R, TB>), dim3(grid), dim3(1 << (TB - 4)), 0, s, \
                                           (const Cx<R>*)src, (Cx<R>*)dst, pa, sa); break;
        SC(10) SC(11) SC(12) SC(13)
#undef SC
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

template <typename R>
int copy_tile_r(int tile_bits, void* state, const PassArgs& pa, hipStream_t s)
{
    const unsigned grid = 1u << pa.nrest;
    switch (tile_bits) {
#define CT(TB) case TB: hipLaunchKernelGGL((k_copy_tile<R, TB>), dim3(grid), dim3(1 << (TB - 4)), 0, s, (Cx<R>*)state, pa, (R)1); break;
        CT(10) CT(11) CT(12) CT(13)
#undef CT
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

inline unsigned grid_for(uint64_t n)
{
    const uint64_t b = (n + 255) / 256;
    return (unsigned)(b < 65536 ? (b ? b : 1) : 65536);
}

}  // namespace

int launch_batch(int dtype, const BatchArgs& a, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    const uint64_t nw = (uint64_t)a.npasses + 2;  // done[] + next + err
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nw)), dim3(256), 0, s, a.done, (uint64_t)a.npasses, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, a.next, (uint64_t)1, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, a.err, (uint64_t)1, 0u);
    if (dtype == 0) {
        hipLaunchKernelGGL(k_batch_init<double>, dim3(grid_for(a.total_amps)), dim3(256), 0, s, (Cx<double>*)a.states,
                           a.total_amps, a.circ_off, a.count);
        hipLaunchKernelGGL(k_batch_ones<double>, dim3((a.count + 255) / 256), dim3(256), 0, s, (Cx<double>*)a.states,
                           a.circ_off, a.count);
        int e = launch_batch_kernel_f64(a, s);
        if (e) return e;
        hipLaunchKernelGGL(k_batch_z<double>, dim3(a.count), dim3(256), 0, s, (const Cx<double>*)a.states, a.circ_off,
                           a.nq, a.nq_int, a.zoff, a.z, a.norms);
    } else {
        hipLaunchKernelGGL(k_batch_init<float>, dim3(grid_for(a.total_amps)), dim3(256), 0, s, (Cx<float>*)a.states,
                           a.total_amps, a.circ_off, a.count);
        hipLaunchKernelGGL(k_batch_ones<float>, dim3((a.count + 255) / 256), dim3(256), 0, s, (Cx<float>*)a.states,
                           a.circ_off, a.count);
        int e = launch_batch_kernel_f32(a, s);
        if (e) return e;
        hipLaunchKernelGGL(k_batch_z<float>, dim3(a.count), dim3(256), 0, s, (const Cx<float>*)a.states, a.circ_off,
                           a.nq, a.nq_int, a.zoff, a.z, a.norms);
    }
    return (int)hipGetLastError();
}

int launch_overlap(int dtype_a, const void* a, int dtype_b, const void* b, int nq_int, double* d_scratch,
                   double* d_out, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (nq_int < 12) return (int)hipErrorInvalidValue;
    const uint64_t chunks = 1ull << (nq_int - 12);
    double* x = chunks == 1 ? d_out : d_scratch;
    double* y = d_scratch + chunks * 4 + 64;
    const dim3 g((unsigned)chunks), t(256);
    unsigned long long* mx = reinterpret_cast<unsigned long long*>(d_out + 4);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t*>(mx), (uint64_t)2, 0u);
    if (dtype_a == 0 && dtype_b == 0)
        hipLaunchKernelGGL((k_overlap0<double, double>), g, t, 0, s, (const Cx<double>*)a, (const Cx<double>*)b, x, mx);
    else if (dtype_a == 0)
        hipLaunchKernelGGL((k_overlap0<double, float>), g, t, 0, s, (const Cx<double>*)a, (const Cx<float>*)b, x, mx);
    else if (dtype_b == 0)
        hipLaunchKernelGGL((k_overlap0<float, double>), g, t, 0, s, (const Cx<float>*)a, (const Cx<double>*)b, x, mx);
    else
        hipLaunchKernelGGL((k_overlap0<float, float>), g, t, 0, s, (const Cx<float>*)a, (const Cx<float>*)b, x, mx);
    uint64_t rows = chunks;
    while (rows > 1) {
        const uint64_t blocks = (rows + 4095) / 4096;
        double* dst = blocks == 1 ? d_out : y;
        hipLaunchKernelGGL(k_rowsum, dim3((unsigned)blocks), dim3(256), 0, s, x, rows, 4, dst);
        rows = blocks;
        std::swap(x, y);
    }
    return (int)hipGetLastError();
}

int launch_pass(int dtype, int tile_bits, void* state, int nq_int, const PassArgs& pa, const DevPlan& dp, void* stream)
{
    if (pa.tile_bits != tile_bits || nq_int < tile_bits || nq_int - tile_bits > 31 || pa.nrest > nq_int - tile_bits)
        return (int)hipErrorInvalidValue;
    if (dtype == 0) return launch_pass_f64(tile_bits, state, nq_int, pa, dp, stream);
    return launch_pass_f32(tile_bits, state, nq_int, pa, dp, stream);
}

int launch_scatter(int dtype, int tile_bits, const void* src, void* dst, const PassArgs& pa, const ScatterArgs* sa,
                   void* stream)
{
    if (!pa.scatter || pa.tile_bits != tile_bits || !sa || !dst) return (int)hipErrorInvalidValue;
    if (dtype == 0) return scatter_r<double>(tile_bits, src, dst, pa, sa, (hipStream_t)stream);
    return scatter_r<float>(tile_bits, src, dst, pa, sa, (hipStream_t)stream);
}

int launch_copy_tiles(int dtype, int tile_bits, void* state, int nq_int, const PassArgs& pa, void* stream)
{
    if (pa.tile_bits != tile_bits || nq_int < tile_bits || nq_int - tile_bits > 31 || pa.nrest > nq_int - tile_bits)
        return (int)hipErrorInvalidValue;
    if (dtype == 0) return copy_tile_r<double>(tile_bits, state, pa, (hipStream_t)stream);
    return copy_tile_r<float>(tile_bits, state, pa, (hipStream_t)stream);
}

int launch_init_basis(int dtype, void* state, int nq_int, uint64_t index, void* stream)
{
    const uint64_t n = 1ull << nq_int;
    if (dtype == 0)
        hipLaunchKernelGGL(k_init_basis<double>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                           (Cx<double>*)state, n, index);
    else
        hipLaunchKernelGGL(k_init_basis<float>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                           (Cx<float>*)state, n, index);
    return (int)hipGetLastError();
}

int launch_init_product(int dtype, void* state, int nq_int, const double* d_factors, void* stream)
{
    const uint64_t n = 1ull << nq_int;
    if (dtype == 0)
        hipLaunchKernelGGL(k_init_product<double>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                           (Cx<double>*)state, n, nq_int, d_factors);
    else
        hipLaunchKernelGGL(k_init_product<float>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                           (Cx<float>*)state, n, nq_int, d_factors);
    return (int)hipGetLastError();
}

int launch_scale(int dtype, void* state, int nq_int, double factor, void* stream)
{
    const uint64_t n = 1ull << nq_int;
    if (dtype == 0)
        hipLaunchKernelGGL(k_scale<double>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (Cx<double>*)state,
                           n, factor);
    else
        hipLaunchKernelGGL(k_scale<float>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (Cx<float>*)state, n,
                           factor);
    return (int)hipGetLastError();
}

uint64_t reduce_scratch_doubles(int nq_int)
{
    // level-0 records (26 doubles per 4096 amplitudes) twice over, plus the
    // Z partials (16 per 4096), with slack for the small levels.
    const uint64_t chunks = 1ull << (nq_int > 12 ? nq_int - 12 : 0);
    return chunks * 26 * 2 + chunks * 16 * 2 + 4096;
}

int launch_amp_sums(int dtype, const void* state, int nq_int, double* d_scratch, double* d_out, void* stream,
                    const uint64_t* gather_cols)
{
    hipStream_t s = (hipStream_t)stream;
    if (nq_int < 12 || nq_int > kMaxQubits) return (int)hipErrorInvalidValue;
    uint64_t count = 1ull << (nq_int - 12);
    double* a = count == 1 ? d_out : d_scratch;
    double* b = d_scratch + count * 26 + 64;
    GatherMap gm{};
    if (gather_cols) {
        const uint64_t lim = nq_int >= 64 ? ~0ull : ((1ull << nq_int) - 1ull);
        for (int q = 0; q < nq_int; q++) {
            if (gather_cols[q] & ~lim) return (int)hipErrorInvalidValue;  // offsets stay inside the state
            gm.col[q] = gather_cols[q];
        }
    }
    if (dtype == 0) {
        if (gather_cols)
            hipLaunchKernelGGL((k_sums0<double, true>), dim3((unsigned)count), dim3(256), 0, s,
                               (const Cx<double>*)state, a, gm);
        else
            hipLaunchKernelGGL((k_sums0<double, false>), dim3((unsigned)count), dim3(256), 0, s,
                               (const Cx<double>*)state, a, gm);
    } else {
        if (gather_cols)
            hipLaunchKernelGGL((k_sums0<float, true>), dim3((unsigned)count), dim3(256), 0, s,
                               (const Cx<float>*)state, a, gm);
        else
            hipLaunchKernelGGL((k_sums0<float, false>), dim3((unsigned)count), dim3(256), 0, s,
                               (const Cx<float>*)state, a, gm);
    }
    int K = 12;
    while (count > 1) {
        int C = 0;
        while (C < 8 && (1ull << (C + 1)) <= count) C++;
        const uint64_t blocks = count >> C;
        double* dst = (blocks == 1) ? d_out : b;
        hipLaunchKernelGGL(k_sums_rec, dim3((unsigned)blocks), dim3(256), 0, s, a, K, C, dst);
        K += C;
        count = blocks;
        std::swap(a, b);
        if (blocks == 1) break;
    }
    return (int)hipGetLastError();
}

int launch_zsums(int dtype, const void* state, int nq_int, const uint64_t* d_masks, int count, double* d_scratch,
                 double* d_out, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (nq_int < 12 || count < 1 || count > 16) return (int)hipErrorInvalidValue;
    const uint64_t chunks = 1ull << (nq_int - 12);
    double* a = chunks == 1 ? d_out : d_scratch;
    double* b = d_scratch + chunks * 16 + 64;
    if (dtype == 0)
        hipLaunchKernelGGL((k_z0<double, 16>), dim3((unsigned)chunks), dim3(256), 0, s, (const Cx<double>*)state,
                           d_masks, count, a);
    else
        hipLaunchKernelGGL((k_z0<float, 16>), dim3((unsigned)chunks), dim3(256), 0, s, (const Cx<float>*)state,
                           d_masks, count, a);
    uint64_t rows = chunks;
    while (rows > 1) {
        const uint64_t blocks = (rows + 4095) / 4096;
        double* dst = blocks == 1 ? d_out : b;
        hipLaunchKernelGGL(k_rowsum, dim3((unsigned)blocks), dim3(256), 0, s, a, rows, 16, dst);
        rows = blocks;
        std::swap(a, b);
        if (blocks == 1) break;
    }
    return (int)hipGetLastError();
}

}  // namespace kate
