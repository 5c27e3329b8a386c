// kernels_dev.hpp -- device code shared by the kernel translation units
// (kernels.hip, pass_f64.hip, pass_f32.hip, batch_f64.hip, batch_f32.hip).
// Each unit instantiates only the templates it launches, so the pass and
// batch kernels of the two precisions compile in parallel.
//
// k_pass is the gate-group kernel: the replacement of Kate's
// executeGroupKernelSharedState + GPUGate::compute
// (HIPcode/simulator.hpp:143-218, HIPcode/GPUGate.hpp:21-203) and, with a
// permuted store, of its multi-GPU qubit swap (HIPcode/simulator.hpp:102-115).
// One workgroup owns one tile of 2^TB amplitudes:
//   1. load: 16 independent coalesced loads per thread (tile bits 0..TB-5
//      come from the thread index, so each wave reads whole runs of the
//      tile's low qubits) into an XOR-swizzled LDS tile;
//   2. micro-steps: each thread pulls the 16 amplitudes spanned by 4 tile
//      bits into registers, applies every gate of the step there (gate
//      coefficients are uniform, read through the scalar cache), writes them
//      back; one barrier per micro-step;
//   3. store: the same 16 coalesced runs, optionally to permuted positions.
// Complex arithmetic follows Kate's Complex class
// (backendagnosticcode/basic_host_types.hpp:43-45); the compiler may contract
// it into FMAs, so results agree with Kate's CPU path to rounding, not bits.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdlib>

#include "kernels.hpp"

namespace kate {
namespace {

template <typename R>
struct alignas(2 * sizeof(R)) Cx {
    R re, im;
};

// Fused multiply-add written out: which product of an expression is fused
// is then fixed by the source, not by how the compiler contracts each
// inlined copy, so the direct and LDS paths of a step round identically.
template <typename R>
__device__ __forceinline__ R fmar(R a, R b, R c)
{
    if constexpr (sizeof(R) == 8) return __builtin_fma(a, b, c);
    else return __builtin_fmaf(a, b, c);
}

template <typename R>
__device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> b)
{
    return Cx<R>{fmar(a.re, b.re, -(a.im * b.im)), fmar(a.re, b.im, a.im * b.re)};
}

template <typename R>
__device__ __forceinline__ Cx<R> cmac(Cx<R> acc, Cx<R> a, Cx<R> b)
{
    acc.re = fmar(-a.im, b.im, fmar(a.re, b.re, acc.re));
    acc.im = fmar(a.im, b.re, fmar(a.re, b.im, acc.im));
    return acc;
}

// XOR swizzle of the low W bits of an LDS element index with the next W bits,
// W = log2(elements per 256 B bank row): makes the 16-amplitude micro-step
// accesses of a wave land in distinct banks for every choice of tile bits
// inside the low 2W bits.
template <typename R>
__device__ __forceinline__ uint32_t swz(uint32_t j)
{
    constexpr uint32_t W = sizeof(R) == 8 ? 4 : 5;
    return j ^ ((j >> W) & ((1u << W) - 1u));
}

// Deterministic reductions: fixed butterfly within a wave, fixed order across waves.
__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Sum over the 256 threads of a block, fixed order; result valid in thread 0.
__device__ __forceinline__ double block_sum(double x, double* lds4)
{
    x = wave_sum(x);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds4[w] = x;
    __syncthreads();
    return ((lds4[0] + lds4[1]) + (lds4[2] + lds4[3]));
}

// Deposit the low bits of s into the set bits of MASK (compile-time).
__host__ __device__ constexpr int dep(int mask, int s)
{
    int out = 0, k = 0;
    for (int b = 0; b < 4; b++)
        if ((mask >> b) & 1) {
            if ((s >> k) & 1) out |= 1 << b;
            k++;
        }
    return out;
}

__host__ __device__ constexpr int popc4(int m) { return (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1) + ((m >> 3) & 1); }

// ---- register ops on the N = 2^NB amplitudes of a micro-step --------------
//
// A micro-step holds N = 8 (NB = 3 micro bits) or N = 16 (NB = 4) amplitudes
// per thread; register index bit i is micro bit i.  Every op is a template
// on the array size, and apply_op instantiates only the codes whose bits
// exist for that size.

// Dense gate on the micro bits of MASK; matrix bit i <-> i-th lowest bit.
template <int MASK, typename R, int N>
__device__ __forceinline__ void op_dense(Cx<R> (&v)[N], const Cx<R>* __restrict__ m)
{
    constexpr int K = popc4(MASK);
    constexpr int D = 1 << K;
#pragma unroll
    for (int rest = 0; rest < N; rest++) {
        if (rest & MASK) continue;
        Cx<R> in[D];
#pragma unroll
        for (int s = 0; s < D; s++) in[s] = v[rest | dep(MASK, s)];
#pragma unroll
        for (int r = 0; r < D; r++) {
            Cx<R> acc = cmul(m[r * D], in[0]);
#pragma unroll
            for (int c = 1; c < D; c++) acc = cmac(acc, m[r * D + c], in[c]);
            v[rest | dep(MASK, r)] = acc;
        }
    }
}

template <int J, typename R, int N>
__device__ __forceinline__ void op_phase(Cx<R> (&v)[N], const Cx<R>* __restrict__ m)
{
    const Cx<R> d0 = m[0], d1 = m[1];
#pragma unroll
    for (int s = 0; s < N; s++) v[s] = cmul(((s >> J) & 1) ? d1 : d0, v[s]);
}

// Kate's Hadamard: a0' = (a0+a1)/sqrt2, a1' = a0' - a1*sqrt2 (GateComputing.hpp:77-78).
template <int J, typename R, int N>
__device__ __forceinline__ void op_h(Cx<R> (&v)[N])
{
    const R inv = (R)0.7071067811865475, s2 = (R)1.4142135623730951;
#pragma unroll
    for (int s = 0; s < N; s++) {
        if ((s >> J) & 1) continue;
        const Cx<R> a = v[s], b = v[s | (1 << J)];
        const Cx<R> n0{(a.re + b.re) * inv, (a.im + b.im) * inv};
        v[s] = n0;
        v[s | (1 << J)] = Cx<R>{fmar(-b.re, s2, n0.re), fmar(-b.im, s2, n0.im)};
    }
}

template <int T, int C, typename R, int N>
__device__ __forceinline__ void op_cnot(Cx<R> (&v)[N])
{
#pragma unroll
    for (int s = 0; s < N; s++) {
        if (((s >> C) & 1) == 0 || ((s >> T) & 1)) continue;
        const Cx<R> t = v[s];
        v[s] = v[s | (1 << T)];
        v[s | (1 << T)] = t;
    }
}

template <int A, int B, typename R, int N>
__device__ __forceinline__ void op_crk(Cx<R> (&v)[N], const Cx<R>* __restrict__ m)
{
    const Cx<R> ph = m[0];
#pragma unroll
    for (int s = 0; s < N; s++)
        if (((s >> A) & 1) && ((s >> B) & 1)) v[s] = cmul(v[s], ph);
}

template <int T, int C0, int C1, typename R, int N>
__device__ __forceinline__ void op_toff(Cx<R> (&v)[N])
{
#pragma unroll
    for (int s = 0; s < N; s++) {
        if (((s >> C0) & 1) == 0 || ((s >> C1) & 1) == 0 || ((s >> T) & 1)) continue;
        const Cx<R> t = v[s];
        v[s] = v[s | (1 << T)];
        v[s | (1 << T)] = t;
    }
}

// Independent 2x2 gates on every micro bit of MASK (one matrix per bit, in
// ascending bit order), applied one after the other.
template <int MASK, typename R, int N>
__device__ __forceinline__ void op_kron(Cx<R> (&v)[N], const Cx<R>* __restrict__ m)
{
    int k = 0;
    if (MASK & 1) op_dense<1>(v, m + 4 * k++);
    if (MASK & 2) op_dense<2>(v, m + 4 * k++);
    if (MASK & 4) op_dense<4>(v, m + 4 * k++);
    if constexpr (N > 8)
        if (MASK & 8) op_dense<8>(v, m + 4 * k++);
}

// Packed index of the bits of s selected by MASK (inverse of dep).
__host__ __device__ constexpr int ext(int mask, int s)
{
    int out = 0, k = 0;
    for (int b = 0; b < 4; b++)
        if ((mask >> b) & 1) {
            if ((s >> b) & 1) out |= 1 << k;
            k++;
        }
    return out;
}

// Kronecker product of 2x2 gates in Euler form on the micro bits of MASK:
// gate b is D1_b * RX_b * D2_b with RX_b = [[x, -iy], [-iy, x]] (x, y real),
// D1_b diagonal and D2_b = diag(1, q_b).  The rotations cost 4 FMA-class
// instructions per amplitude each (against 8 for a dense 2x2), and the
// diagonals of all bits are one multiply by a table entry indexed by the
// amplitude's MASK bits.  Coefficients (complex elements): (x_b, y_b) per
// bit in ascending order, the D1 table (2^K entries), then, with D2, the D2
// table (its entry 0 is 1 and is not applied).
template <int MASK, bool HAS_D2, typename R, int N>
__device__ __forceinline__ void op_ekron(Cx<R> (&v)[N], const Cx<R>* __restrict__ m)
{
    constexpr int K = popc4(MASK);
    if constexpr (HAS_D2) {
        const Cx<R>* __restrict__ t2 = m + K + (1 << K);
#pragma unroll
        for (int s = 0; s < N; s++)
            if (ext(MASK, s)) v[s] = cmul(t2[ext(MASK, s)], v[s]);
    }
    int k = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        if (!((MASK >> b) & 1)) continue;
        const R x = m[k].re, y = m[k].im;
        k++;
#pragma unroll
        for (int s = 0; s < N; s++) {
            if ((s >> b) & 1) continue;
            const Cx<R> a = v[s], c = v[s | (1 << b)];
            v[s] = Cx<R>{fmar(x, a.re, y * c.im), fmar(x, a.im, -(y * c.re))};
            v[s | (1 << b)] = Cx<R>{fmar(x, c.re, y * a.im), fmar(x, c.im, -(y * a.re))};
        }
    }
    const Cx<R>* __restrict__ t1 = m + K;
#pragma unroll
    for (int s = 0; s < N; s++) v[s] = cmul(t1[ext(MASK, s)], v[s]);
}

// OK_FKRON on every register bit: the rotations scaled to
// [[1, -it], [-it, 1]] cost 2 FMAs per amplitude per bit (one per output
// component), then one multiply by the table T (the scales, the D1
// diagonal); coefficients: (t_b, 0) per bit in ascending order, then T.
// The exchanged outputs of the flip mask are undone by the caller's store.
template <typename R, int N>
__device__ __forceinline__ void op_fkron(Cx<R> (&v)[N], const Cx<R>* __restrict__ m)
{
#pragma unroll
    for (int b = 0; (1 << b) < N; b++) {
        const R t = m[b].re;
#pragma unroll
        for (int s = 0; s < N; s++) {
            if ((s >> b) & 1) continue;
            const Cx<R> a = v[s], c = v[s | (1 << b)];
            v[s] = Cx<R>{fmar(t, c.im, a.re), fmar(-t, c.re, a.im)};
            v[s | (1 << b)] = Cx<R>{fmar(t, a.im, c.re), fmar(-t, a.re, c.im)};
        }
    }
    constexpr int K = N == 16 ? 4 : (N == 8 ? 3 : 2);
    const Cx<R>* __restrict__ t1 = m + K;
#pragma unroll
    for (int s = 0; s < N; s++) v[s] = cmul(t1[s], v[s]);
}

#define KATE_OPC(kind, var) (((kind) << 8) | (var))

template <typename R, int N>
__device__ __forceinline__ void apply_op(Cx<R> (&v)[N], uint32_t code, const Cx<R>* __restrict__ cf)
{
    switch (code) {
#define D(M) \
    case KATE_OPC(OK_DENSE, M): if constexpr ((M) < N) op_dense<M>(v, cf); break; \
    case KATE_OPC(OK_KRON, M): if constexpr ((M) < N) op_kron<M>(v, cf); break;
        D(1) D(2) D(3) D(4) D(5) D(6) D(7) D(8) D(9) D(10) D(11) D(12) D(13) D(14) D(15)
#undef D
#define E(M) \
    case KATE_OPC(OK_EKRON, M): if constexpr ((M) < N) op_ekron<M, false>(v, cf); break; \
    case KATE_OPC(OK_EKRON, (M) | 16): if constexpr ((M) < N) op_ekron<M, true>(v, cf); break;
        E(3) E(5) E(6) E(7) E(9) E(10) E(11) E(12) E(13) E(14) E(15)
#undef E
#define P(J) \
    case KATE_OPC(OK_PHASE, J): if constexpr ((1 << (J)) < N) op_phase<J>(v, cf); break; \
    case KATE_OPC(OK_H, J): if constexpr ((1 << (J)) < N) op_h<J>(v); break;
        P(0) P(1) P(2) P(3)
#undef P
#define CN(T, C) \
    case KATE_OPC(OK_CNOT, (T) * 4 + (C)): if constexpr ((1 << (T)) < N && (1 << (C)) < N) op_cnot<T, C>(v); break;
        CN(0, 1) CN(0, 2) CN(0, 3) CN(1, 0) CN(1, 2) CN(1, 3) CN(2, 0) CN(2, 1) CN(2, 3) CN(3, 0) CN(3, 1) CN(3, 2)
#undef CN
#define CR(A, B) \
    case KATE_OPC(OK_CRK, (A) * 4 + (B)): if constexpr ((1 << (A)) < N && (1 << (B)) < N) op_crk<A, B>(v, cf); break;
        CR(0, 1) CR(0, 2) CR(0, 3) CR(1, 2) CR(1, 3) CR(2, 3)
#undef CR
#define TF(T, A, B) \
    case KATE_OPC(OK_TOFF, (T) * 16 + (A) * 4 + (B)): \
        if constexpr ((1 << (T)) < N && (1 << (A)) < N && (1 << (B)) < N) op_toff<T, A, B>(v); \
        break;
        TF(0, 1, 2) TF(0, 1, 3) TF(0, 2, 3) TF(1, 0, 2) TF(1, 0, 3) TF(1, 2, 3)
        TF(2, 0, 1) TF(2, 0, 3) TF(2, 1, 3) TF(3, 0, 1) TF(3, 0, 2) TF(3, 1, 2)
#undef TF
    default:
        break;
    }
}

// The op set of passes the planner marks light (PassArgs::light): Euler-form
// Kronecker ops without a right diagonal, dense 1-qubit ops, phases,
// Hadamards and CNOTs -- what RX/RZ layers with CNOT ladders lower to.  A
// shorter decision tree than apply_op, and no multi-qubit dense op holding
// its inputs and outputs at once.
template <typename R, int N>
__device__ __forceinline__ void apply_op_light(Cx<R> (&v)[N], uint32_t code, const Cx<R>* __restrict__ cf)
{
    switch (code) {
#define D1(M) \
    case KATE_OPC(OK_DENSE, M): if constexpr ((M) < N) op_dense<M>(v, cf); break;
        D1(1) D1(2) D1(4) D1(8)
#undef D1
#define E(M) \
    case KATE_OPC(OK_EKRON, M): if constexpr ((M) < N) op_ekron<M, false>(v, cf); break;
        E(3) E(5) E(6) E(7) E(9) E(10) E(11) E(12) E(13) E(14) E(15)
#undef E
#define P(J) \
    case KATE_OPC(OK_PHASE, J): if constexpr ((1 << (J)) < N) op_phase<J>(v, cf); break; \
    case KATE_OPC(OK_H, J): if constexpr ((1 << (J)) < N) op_h<J>(v); break;
        P(0) P(1) P(2) P(3)
#undef P
#define CN(T, C) \
    case KATE_OPC(OK_CNOT, (T) * 4 + (C)): if constexpr ((1 << (T)) < N && (1 << (C)) < N) op_cnot<T, C>(v); break;
        CN(0, 1) CN(0, 2) CN(0, 3) CN(1, 0) CN(1, 2) CN(1, 3) CN(2, 0) CN(2, 1) CN(2, 3) CN(3, 0) CN(3, 1) CN(3, 2)
#undef CN
    default:
        break;
    }
}

// LDS index of tile index y under a layout (see MicroStep::cols).
template <int TB>
__device__ __forceinline__ uint32_t layout(uint32_t y, const uint16_t* __restrict__ cols)
{
    uint32_t p = 0;
#pragma unroll
    for (int b = 0; b < TB; b++)
        if ((y >> b) & 1u) p ^= cols[b];
    return p;
}

// One dense gate of any width on the LDS tile: every thread computes its
// 2^NB outputs from LDS, then (after a barrier) writes them.
template <typename R, int TB, int NB>
__device__ __forceinline__ void generic_dense(Cx<R>* tile, const MicroStep* __restrict__ msp,
                                              const Cx<R>* __restrict__ m)
{
    const MicroStep& ms = *msp;
    constexpr int NT = 1 << (TB - NB);
    constexpr int N = 1 << NB;
    const int k = ms.nbits;
    const uint32_t dim = 1u << k;
    uint32_t gmask = 0;
    for (int i = 0; i < k; i++) gmask |= 1u << ms.bits[i];
    Cx<R> out[N];
#pragma unroll
    for (int e = 0; e < N; e++) {
        const uint32_t j = (uint32_t)e * NT + threadIdx.x;
        uint32_t row = 0;
        for (int i = 0; i < k; i++) row |= ((j >> ms.bits[i]) & 1u) << i;
        const uint32_t base = layout<TB>(j & ~gmask, ms.cols);
        Cx<R> acc{0, 0};
        for (uint32_t c = 0; c < dim; c++) {
            uint32_t idx = base;
            for (int i = 0; i < k; i++)
                if ((c >> i) & 1u) idx ^= ms.cols[ms.bits[i]];
            acc = cmac(acc, m[row * dim + c], tile[swz<R>(idx)]);
        }
        out[e] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < N; e++) tile[swz<R>(layout<TB>((uint32_t)e * NT + threadIdx.x, ms.cols))] = out[e];
}

// One dense 5-qubit gate on the LDS tile (an SK_GENERIC step of 5 bits):
// the tile splits into 2^(TB-5) groups of 32 amplitudes that the gate mixes.
// Lane l of a wave holds group l (mod the group count) and all threads of a
// wave compute the same rows, so the matrix entries are wave-uniform and
// come in through scalar loads, by columns from the transpose that the plan
// stores after the matrix.  Each thread streams its group's 32 amplitudes
// from LDS into its rows' accumulators and, after a barrier, writes them.
// Every row sums its products in generic_dense's order, so the results are
// bit for bit the same.
template <typename R, int TB, int NB>
__device__ __forceinline__ void dense5(Cx<R>* tile, const MicroStep* __restrict__ msp,
                                       const Cx<R>* __restrict__ m)
{
    constexpr int NT = 1 << (TB - NB);  // threads of the workgroup
    constexpr int G = 1 << (TB - 5);    // groups of 32 amplitudes
    constexpr int RPT = 32 * G / NT;    // rows per thread
    static_assert(NT % G == 0 && RPT >= 1, "each group needs a whole number of threads");
    const uint32_t tid = threadIdx.x;
    const uint32_t grp = tid % (uint32_t)G;
    uint32_t part = tid / (uint32_t)G;
    if constexpr (G >= 64) part = __builtin_amdgcn_readfirstlane(part);  // one part per wave
    uint32_t gmask = 0, oc[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        gmask |= 1u << msp->bits[i];
        oc[i] = msp->cols[msp->bits[i]];
    }
    // the group's tile index: its bits on the non-gate tile bits, ascending
    uint32_t y = 0, g = grp;
#pragma unroll
    for (int b = 0; b < TB; b++)
        if (!((gmask >> b) & 1u)) {
            y |= (g & 1u) << b;
            g >>= 1;
        }
    const uint32_t base = layout<TB>(y, msp->cols);
    const uint32_t row0 = part * (uint32_t)RPT;
    const Cx<R>* __restrict__ mt = m + 1024 + row0;  // transpose: entry (row, c) at c * 32 + row
    Cx<R> acc[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) acc[r] = Cx<R>{0, 0};
#pragma unroll 4
    for (int c = 0; c < 32; c++) {
        uint32_t idx = base;
#pragma unroll
        for (int i = 0; i < 5; i++)
            if ((c >> i) & 1) idx ^= oc[i];
        const Cx<R> x = tile[swz<R>(idx)];
#pragma unroll
        for (int r = 0; r < RPT; r++) acc[r] = cmac(acc[r], mt[c * 32 + r], x);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const uint32_t row = row0 + (uint32_t)r;
        uint32_t idx = base;
#pragma unroll
        for (int i = 0; i < 5; i++)
            if ((row >> i) & 1u) idx ^= oc[i];
        tile[swz<R>(idx)] = acc[r];
    }
}

// WIDE instantiations also run SK_GENERIC micro-steps; keeping them out of
// the common variant keeps its register count at two waves per SIMD.
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Dense K-qubit gate (K = 4 or 5) on the matrix cores, complex64 tile.
// The gate is a real 2D x 2D matrix (D = 2^K) in the embedding
// [Yr; Yi] = [[Mr, -Mi], [Mi, Mr]] [Xr; Xi], with the contraction index
// ordered (gate row g, re/im) so each lane reads one float of X per step.
// The tile is the 2D x 2^(TB-K) right-hand side; each wave owns whole
// 32-column blocks, so it reads and overwrites only its own columns.
// A is pre-arranged on the host in lane order: A[(rb*D + s)*64 + lane] =
// A'[rb*32 + lane%32][2s + lane/32] for v_mfma_f32_32x32x2f32.
template <int TB, int K>
__device__ __forceinline__ void mfma_dense(Cx<float>* tile, const MicroStep* __restrict__ msp,
                                           const float* __restrict__ A)
{
    constexpr int D = 1 << K;
    constexpr int RB = 2 * D / 32;
    constexpr int NW = (1 << (TB - 4)) / 64;
    constexpr int CB = (1 << (TB - K)) / 32;
    constexpr int CBW = CB / NW;
    static_assert(CBW >= 1 && CBW * NW == CB, "column blocks must split evenly over the waves");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t half = (uint32_t)lane >> 5;
    float* tf = reinterpret_cast<float*>(tile);
    float a[RB][D];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int s = 0; s < D; s++) a[rb][s] = A[(rb * D + s) * 64 + lane];
    uint32_t gc[K];
#pragma unroll
    for (int i = 0; i < K; i++) gc[i] = msp->gcols[i];
#pragma unroll
    for (int cbi = 0; cbi < CBW; cbi++) {
        const uint32_t col = (uint32_t)(wave * CBW + cbi) * 32u + (uint32_t)(lane & 31);
        uint32_t lq = 0;
#pragma unroll
        for (int j = 0; j < TB - K; j++)
            if ((col >> j) & 1u) lq ^= msp->tcols[j];
        f32x16 acc[RB];
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[rb][r] = 0.0f;
#pragma unroll
        for (int s = 0; s < D; s++) {
            uint32_t lp = lq;
#pragma unroll
            for (int i = 0; i < K; i++)
                if ((s >> i) & 1) lp ^= gc[i];
            const float b = tf[2 * swz<float>(lp) + half];
#pragma unroll
            for (int rb = 0; rb < RB; rb++) acc[rb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[rb][s], b, acc[rb], 0, 0, 0);
        }
        // acc[rb][r] holds output row rb*32 + 8(r/4) + 4*half + r%4 of this lane's column
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int o0 = rb * 32 + 8 * (r / 4) + (r % 4);  // row without the lane-half bit (bit 2)
                const int g0 = o0 % D;
                uint32_t lp = lq;
#pragma unroll
                for (int i = 0; i < K; i++)
                    if (i != 2 && ((g0 >> i) & 1)) lp ^= gc[i];
                if (half) lp ^= gc[2];
                tf[2 * swz<float>(lp) + (uint32_t)(o0 / D)] = acc[rb][r];
            }
    }
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// mfma_dense with split-bf16 products: A = Ahi + Alo (split on the host),
// B = Bhi + Blo (split here), acc += Ahi Bhi + Ahi Blo + Alo Bhi on
// v_mfma_f32_32x32x16_bf16 (lane l holds contraction indices 8(l/32)..+7 of
// its row / column).  Eight consecutive contraction indices are four complex
// amplitudes of consecutive gate rows, read as four 8-byte LDS loads.
template <int TB, int K>
__device__ __forceinline__ void mfma_bf16(Cx<float>* tile, const MicroStep* __restrict__ msp,
                                          const uint32_t* __restrict__ A)
{
    constexpr int D = 1 << K;
    constexpr int RB = 2 * D / 32;
    constexpr int T = 2 * D / 16;
    constexpr int NW = (1 << (TB - 4)) / 64;
    constexpr int CB = (1 << (TB - K)) / 32;
    constexpr int CBW = CB / NW;
    static_assert(CBW >= 1 && CBW * NW == CB, "column blocks must split evenly over the waves");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t half = (uint32_t)lane >> 5;
    u32x4 a[RB][T][2];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int t = 0; t < T; t++)
#pragma unroll
            for (int part = 0; part < 2; part++)
                a[rb][t][part] = *reinterpret_cast<const u32x4*>(A + ((size_t)((rb * T + t) * 2 + part) * 64 + lane) * 4);
    uint32_t gc[K];
#pragma unroll
    for (int i = 0; i < K; i++) gc[i] = msp->gcols[i];
    const uint32_t hbit = half ? gc[2] : 0u;
#pragma unroll
    for (int cbi = 0; cbi < CBW; cbi++) {
        const uint32_t col = (uint32_t)(wave * CBW + cbi) * 32u + (uint32_t)(lane & 31);
        uint32_t lq = 0;
#pragma unroll
        for (int j = 0; j < TB - K; j++)
            if ((col >> j) & 1u) lq ^= msp->tcols[j];
        f32x16 acc[RB];
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[rb][r] = 0.0f;
#pragma unroll
        for (int t = 0; t < T; t++) {
            uint32_t hi[4], lo[4];
#pragma unroll
            for (int jj = 0; jj < 4; jj++) {
                const int g = 8 * t + jj;  // plus 4*half
                uint32_t lp = lq ^ hbit;
#pragma unroll
                for (int i = 0; i < K; i++)
                    if (i != 2 && ((g >> i) & 1)) lp ^= gc[i];
                const Cx<float> x = tile[swz<float>(lp)];
                // x = hi + lo in bf16, both rounded to nearest even by the
                // packed conversion (re in the low half, im in the high half)
                const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{x.re, x.im}, bf16x2));
                const f32x2 rest{x.re - __uint_as_float(h << 16), x.im - __uint_as_float(h & 0xFFFF0000u)};
                hi[jj] = h;
                lo[jj] = __builtin_bit_cast(uint32_t, __builtin_convertvector(rest, bf16x2));
            }
            const bf16x8 bh = __builtin_bit_cast(bf16x8, u32x4{hi[0], hi[1], hi[2], hi[3]});
            const bf16x8 bl = __builtin_bit_cast(bf16x8, u32x4{lo[0], lo[1], lo[2], lo[3]});
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                const bf16x8 ah = __builtin_bit_cast(bf16x8, a[rb][t][0]);
                const bf16x8 al = __builtin_bit_cast(bf16x8, a[rb][t][1]);
                acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[rb], 0, 0, 0);
                acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[rb], 0, 0, 0);
                acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[rb], 0, 0, 0);
            }
        }
        float* tf = reinterpret_cast<float*>(tile);
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int o0 = rb * 32 + 8 * (r / 4) + (r % 4);
                const int g0 = o0 % D;
                uint32_t lp = lq ^ hbit;
#pragma unroll
                for (int i = 0; i < K; i++)
                    if (i != 2 && ((g0 >> i) & 1)) lp ^= gc[i];
                tf[2 * swz<float>(lp) + (uint32_t)(o0 / D)] = acc[rb][r];
            }
    }
}

// State offset of the first amplitude of tile `blk`: its bits go to the
// non-tile qubits rest_pos[0..nrest).  Unrolled with constant indices so that
// every rest_pos[] read is a scalar load and the whole sum stays in SALU;
// a loop indexed by the run-time nrest turns each read into a dependent
// vector load.
__device__ __forceinline__ uint64_t rest_base(const PassArgs& pa, uint64_t blk)
{
    uint64_t base = 0;
    const int nrest = pa.nrest;
#pragma unroll
    for (int i = 0; i < kMaxQubits - 10; i++)
        if (i < nrest) base |= ((blk >> i) & 1ull) << pa.rest_pos[i];
    return base;
}

static_assert(offsetof(MicroStep, op0_code) == 2 && offsetof(MicroStep, coef) == 12 &&
                  offsetof(MicroStep, op_begin) == 4 && offsetof(MicroStep, op_count) == 8 &&
                  offsetof(MicroStep, tbyte) == 112 && offsetof(MicroStep, pbyte) == 160,
              "step_addressing reads MicroStep as 32-bit words");

// Addressing of an SK_REG micro-step for this thread: op range, the LDS
// byte offsets o[] of the register bits and the thread's base offset b (all
// already swizzled: the swizzle is linear, so offsets combine by XOR).  The
// descriptor is read as 32-bit words at constant offsets (scalar loads).
template <int TB, int NB>
__device__ __forceinline__ void step_addressing(const MicroStep* __restrict__ msp, uint32_t tid, uint32_t& op_begin,
                                                uint32_t& op_count, uint32_t (&o)[4], uint32_t& b)
{
    const uint32_t* __restrict__ w = reinterpret_cast<const uint32_t*>(msp);
    op_begin = w[1];
    op_count = w[2];
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = w[40 + k];
    b = 0;
#pragma unroll
    for (int i = 0; i < TB - NB; i++)
        if ((tid >> i) & 1u) b ^= w[28 + i];
}

template <typename R>
__device__ __forceinline__ Cx<R>& lds_at(Cx<R>* tile, uint32_t byte_off)
{
    return *reinterpret_cast<Cx<R>*>(reinterpret_cast<char*>(tile) + byte_off);
}

// LDS offset of register amplitude s from the columns of the micro bits.
template <int NB>
__device__ __forceinline__ uint32_t reg_offset(int s, const uint32_t (&o)[4])
{
    uint32_t x = 0;
#pragma unroll
    for (int k = 0; k < NB; k++)
        if ((s >> k) & 1) x ^= o[k];
    return x;
}

// The micro-steps of one pass on the LDS tile (each thread holds 2^NB
// amplitudes of a step in registers); one barrier after each step.
// The ops of an SK_REG step on the thread's registers.  The first op comes
// with the step (no load that waits on the step); the others are read as
// 32-bit words: a 16-bit field would be a vector load (there are no 16-bit
// scalar loads), whose wait also drains every other vector memory operation
// of the wave.
// Returns the step's flip mask: the store writes register s where register
// s ^ flip belongs (non-zero only for a lone OK_FKRON op).
template <typename R, int N, bool LIGHT>
__device__ __forceinline__ uint32_t reg_step_ops(Cx<R> (&v)[N], const MicroStep* __restrict__ msp, uint32_t op_begin,
                                                 uint32_t op_count, const MicroOp* __restrict__ ops,
                                                 const Cx<R>* __restrict__ coef)
{
    const uint32_t w0 = reinterpret_cast<const uint32_t*>(msp)[0];
    const uint32_t code0 = w0 >> 16, coef0 = reinterpret_cast<const uint32_t*>(msp)[3];
    // the planner makes OK_FKRON only in light passes, as the lone op of a
    // step, on every register bit
    if constexpr (LIGHT) {
        if (op_count == 1 && (code0 >> 8) == OK_FKRON) {
            op_fkron<R, N>(v, coef + coef0);
            return (code0 >> 4) & (N - 1);
        }
    }
    // Light complex128 passes: the common step (RX/RZ layers), one Kronecker
    // op in Euler form on every register bit, is taken before the op switch
    // (in the complex64 kernels the extra path would cost occupancy).
    if constexpr (LIGHT && sizeof(R) == 8) {
        if (op_count == 1 && code0 == KATE_OPC(OK_EKRON, N - 1)) {
            op_ekron<N - 1, false>(v, coef + coef0);
            return 0;
        }
    }
    for (uint32_t o = 0; o < op_count; o++) {
        uint32_t code = code0, cofs = coef0;
        if (o) {
            const uint32_t* ow = reinterpret_cast<const uint32_t*>(ops + op_begin + o);
            code = ow[0] & 0xFFFFu;
            cofs = ow[1];
        }
        if constexpr (LIGHT) apply_op_light<R>(v, code, coef + cofs);
        else apply_op<R>(v, code, coef + cofs);
    }
    return 0;
}

// Amplitude at state offset `off` of the basis state that the first pass
// of an execute makes in registers instead of reading (PassArgs::synth_basis).
template <typename R>
__device__ __forceinline__ Cx<R> synth_amp(const PassArgs& pa, uint64_t off)
{
    return Cx<R>{off == pa.synth_index ? (R)1 : (R)0, (R)0};
}

// Amplitude at `off` in a run from |0...0>: zero without a read when the
// offset has a bit of a qubit that no gate has touched yet.
template <typename R>
__device__ __forceinline__ Cx<R> masked_load(const Cx<R>* __restrict__ state, const PassArgs& pa, uint64_t off)
{
    return (off & pa.zero_mask) ? Cx<R>{(R)0, (R)0} : state[off];
}

// State offset of register amplitude s of this thread for a direct step:
// the XOR of the columns of the set thread-index and register bits.
template <int TB, int NB>
__device__ __forceinline__ uint64_t direct_thread_offset(const uint64_t* __restrict__ g, uint32_t tid)
{
    uint64_t off = 0;
#pragma unroll
    for (int i = 0; i < TB - NB; i++)
        if ((tid >> i) & 1u) off ^= g[i];
    return off;
}

template <int NB>
__device__ __forceinline__ uint64_t direct_reg_offset(const uint64_t* __restrict__ g, int s)
{
    uint64_t off = 0;
#pragma unroll
    for (int k = 0; k < NB; k++)
        if ((s >> k) & 1) off ^= g[12 + k];
    return off;
}

template <typename R, int TB, bool WIDE, int NB, bool LIGHT>
__device__ __forceinline__ void run_steps(const PassArgs& pa, Cx<R>* tile, const MicroStep* __restrict__ micro,
                                          const MicroOp* __restrict__ ops, const Cx<R>* __restrict__ coef,
                                          const uint32_t* __restrict__ raw, uint32_t m_begin, uint32_t m_end)
{
    constexpr int N = 1 << NB;
    const uint32_t tid = threadIdx.x;
    for (uint32_t mi = m_begin; mi < m_end; mi++) {
        const MicroStep* __restrict__ msp = micro + pa.micro_begin + mi;
        if ((reinterpret_cast<const uint32_t*>(msp)[0] & 0xFFu) == SK_REG) {
            uint32_t op_begin, op_count, oc[4], b;
            step_addressing<TB, NB>(msp, tid, op_begin, op_count, oc, b);
            Cx<R> v[N];
#pragma unroll
            for (int s = 0; s < N; s++) v[s] = lds_at(tile, b ^ reg_offset<NB>(s, oc));
            const uint32_t f = reg_step_ops<R, N, LIGHT>(v, msp, op_begin, op_count, ops, coef);
            const uint32_t bw = b ^ reg_offset<NB>((int)f, oc);
#pragma unroll
            for (int s = 0; s < N; s++) lds_at(tile, bw ^ reg_offset<NB>(s, oc)) = v[s];
        } else if (WIDE) {
            // (a 1024-thread workgroup is held to 128 VGPRs: too few for
            // dense5's 32 amplitudes in complex128)
            constexpr bool kDense5 = sizeof(R) == 4 || (TB - NB) <= 9;
            if (kDense5 && msp->kind == SK_GENERIC && msp->nbits == 5) {
                dense5<R, TB, NB>(tile, msp, coef + msp->coef);
            } else if constexpr (sizeof(R) == 4 && NB == 4) {
                Cx<float>* tf = reinterpret_cast<Cx<float>*>(tile);
                if (msp->kind == SK_MFMA) {
                    const float* A = reinterpret_cast<const float*>(raw + msp->coef);
                    if (msp->nbits == 5) mfma_dense<TB, 5>(tf, msp, A);
                    else mfma_dense<TB, 4>(tf, msp, A);
                } else if (msp->kind == SK_MFMA_BF16) {
                    if (msp->nbits == 5) mfma_bf16<TB, 5>(tf, msp, raw + msp->coef);
                    else mfma_bf16<TB, 4>(tf, msp, raw + msp->coef);
                } else {
                    generic_dense<R, TB, NB>(tile, msp, coef + msp->coef);
                }
            } else {
                generic_dense<R, TB, NB>(tile, msp, coef + msp->coef);
            }
        }
        __syncthreads();
    }

}

// Store of one tile, reading it through its final layout: the same
// coalesced runs as the load, optionally to permuted positions (store_pos).
template <typename R, int TB, int NB>
__device__ __forceinline__ void store_tile(Cx<R>* __restrict__ state, const PassArgs& pa, uint64_t base,
                                           Cx<R>* tile)
{
    constexpr int N = 1 << NB;
    const uint32_t tid = threadIdx.x;
    {
        uint64_t tb = base;
        uint32_t lt = 0;
#pragma unroll
        for (int b = 0; b < TB - NB; b++) {
            tb |= (uint64_t)((tid >> b) & 1u) << pa.store_pos[b];
            if ((tid >> b) & 1u) lt ^= pa.store_byte[b];
        }
        uint64_t m[NB];
        uint32_t l[4] = {0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < NB; k++) {
            m[k] = 1ull << pa.store_pos[TB - NB + k];
            l[k] = pa.store_byte[TB - NB + k];
        }
#pragma unroll
        for (int i = 0; i < N; i++) {
            uint64_t off = tb;
#pragma unroll
            for (int k = 0; k < NB; k++)
                if ((i >> k) & 1) off += m[k];
            state[off] = lds_at(tile, lt ^ reg_offset<NB>(i, l));
        }
    }
}

// One tile of one pass: load, micro-steps, store.  `blk` is the tile index
// (its bits are the non-tile qubits rest_pos[]).  Each of the 2^(TB-NB)
// threads holds N = 2^NB amplitudes; the register bits are tile bits
// TB-NB..TB-1 when loading and storing.
// Store of a gate pass that also applies the trailing permutation of its
// circuit (PassArgs::scatter): store index u (thread bits, then register
// bits) reads LDS at the XOR of store_byte[k] over its set bits and writes
// `dst` at the XOR of out_tile[k] and of out_rest[i] over the set bits of
// the tile number; store index bits 0..low-1 are offset bits 0..low-1, so
// the writes are whole runs.
// Output offset of tile `blk` of an out-of-place pass: the XOR of
// out_rest[i] over the set bits of blk (scalar loads, SALU).
__device__ __forceinline__ uint64_t out_base(const PassArgs& pa, uint64_t blk, const ScatterArgs* __restrict__ sa)
{
    uint64_t ob = 0;
#pragma unroll
    for (int i = 0; i < kMaxQubits - 10; i++)
        if (i < pa.nrest && ((blk >> i) & 1ull)) ob ^= sa->out_rest[i];
    return ob;
}

template <typename R, int TB, int NB>
__device__ __forceinline__ void scatter_store(Cx<R>* __restrict__ dst, const PassArgs& pa, uint64_t blk, Cx<R>* tile,
                                              const ScatterArgs* __restrict__ sa)
{
    constexpr int N = 1 << NB;
    const uint32_t tid = threadIdx.x;
    sa += pa.scatter_idx;
    uint64_t ob = out_base(pa, blk, sa);
    uint32_t lt = 0;
#pragma unroll
    for (int b = 0; b < TB - NB; b++)
        if ((tid >> b) & 1u) {
            lt ^= pa.store_byte[b];
            ob ^= sa->out_tile[b];
        }
#pragma unroll
    for (int i = 0; i < N; i++) {
        uint32_t l = lt;
        uint64_t o = ob;
#pragma unroll
        for (int k = 0; k < NB; k++)
            if ((i >> k) & 1) {
                l ^= pa.store_byte[TB - NB + k];
                o ^= sa->out_tile[TB - NB + k];
            }
        dst[o] = lds_at(tile, l);
    }
}

template <typename R, int TB, bool WIDE, int NB, bool LIGHT>
__device__ __forceinline__ void tile_pass(Cx<R>* __restrict__ state, const PassArgs& pa, uint64_t blk, Cx<R>* tile,
                                          const MicroStep* __restrict__ micro, const MicroOp* __restrict__ ops,
                                          const Cx<R>* __restrict__ coef, const uint32_t* __restrict__ raw,
                                          Cx<R>* __restrict__ dst, const ScatterArgs* __restrict__ sa)
{
    constexpr int NT = 1 << (TB - NB);
    constexpr int N = 1 << NB;
    const uint32_t tid = threadIdx.x;

    const uint64_t base = rest_base(pa, blk);
    uint32_t m_begin = 0, m_end = pa.micro_count;

    if (pa.fuse_first) {
        // direct first step: HBM -> registers -> ops -> LDS in the step's layout
        const MicroStep* __restrict__ msp = micro + pa.micro_begin;
        uint32_t op_begin, op_count, oc[4], b;
        step_addressing<TB, NB>(msp, tid, op_begin, op_count, oc, b);
        const uint64_t gt = base ^ direct_thread_offset<TB, NB>(pa.first_g, tid);
        Cx<R> v[N];
        if (pa.synth_basis) {
#pragma unroll
            for (int s = 0; s < N; s++) v[s] = synth_amp<R>(pa, gt ^ direct_reg_offset<NB>(pa.first_g, s));
        } else if (pa.zero_mask) {
#pragma unroll
            for (int s = 0; s < N; s++) v[s] = masked_load(state, pa, gt ^ direct_reg_offset<NB>(pa.first_g, s));
        } else {
#pragma unroll
            for (int s = 0; s < N; s++) v[s] = state[gt ^ direct_reg_offset<NB>(pa.first_g, s)];
        }
        const uint32_t f = reg_step_ops<R, N, LIGHT>(v, msp, op_begin, op_count, ops, coef);
        m_begin = 1;
        if (pa.micro_count == 1 && pa.fuse_last) {
            // the store of one thread can hit an amplitude that another
            // thread of the tile loads (the final layout moves amplitudes
            // between threads): every load of the tile has completed (its
            // value was used above) once all threads pass this barrier
            __syncthreads();
            Cx<R>* __restrict__ out = pa.scatter ? dst : state;
            const uint64_t ob = pa.scatter ? out_base(pa, blk, sa + pa.scatter_idx) : base;
            const uint64_t gs = ob ^ direct_thread_offset<TB, NB>(pa.last_g, tid) ^
                                direct_reg_offset<NB>(pa.last_g, (int)f);
#pragma unroll
            for (int s = 0; s < N; s++) out[gs ^ direct_reg_offset<NB>(pa.last_g, s)] = v[s];
            return;
        }
        const uint32_t bw = b ^ reg_offset<NB>((int)f, oc);
#pragma unroll
        for (int s = 0; s < N; s++) lds_at(tile, bw ^ reg_offset<NB>(s, oc)) = v[s];
    } else {   // load
        uint64_t tb = base;
#pragma unroll
        for (int b = 0; b < TB - NB; b++) tb |= (uint64_t)((tid >> b) & 1u) << pa.load_pos[b];
        uint64_t m[NB];
#pragma unroll
        for (int k = 0; k < NB; k++) m[k] = 1ull << pa.load_pos[TB - NB + k];
        Cx<R> r[N];
        if (pa.synth_basis) {
#pragma unroll
            for (int i = 0; i < N; i++) {
                uint64_t off = tb;
#pragma unroll
                for (int k = 0; k < NB; k++)
                    if ((i >> k) & 1) off += m[k];
                r[i] = synth_amp<R>(pa, off);
            }
        } else if (pa.zero_mask) {
#pragma unroll
            for (int i = 0; i < N; i++) {
                uint64_t off = tb;
#pragma unroll
                for (int k = 0; k < NB; k++)
                    if ((i >> k) & 1) off += m[k];
                r[i] = masked_load(state, pa, off);
            }
        } else {
#pragma unroll
            for (int i = 0; i < N; i++) {
                uint64_t off = tb;
#pragma unroll
                for (int k = 0; k < NB; k++)
                    if ((i >> k) & 1) off += m[k];
                r[i] = state[off];
            }
        }
        const uint32_t tsw = swz<R>(tid) * (uint32_t)sizeof(Cx<R>);
#pragma unroll
        for (int i = 0; i < N; i++) lds_at(tile, tsw ^ (swz<R>((uint32_t)i * NT) * (uint32_t)sizeof(Cx<R>))) = r[i];
    }
    __syncthreads();

    if (pa.fuse_last) m_end--;
    run_steps<R, TB, WIDE, NB, LIGHT>(pa, tile, micro, ops, coef, raw, m_begin, m_end);
    if (pa.fuse_last) {
        // direct last step: LDS in the step's layout -> registers -> ops -> HBM
        const MicroStep* __restrict__ msp = micro + pa.micro_begin + m_end;
        uint32_t op_begin, op_count, oc[4], b;
        step_addressing<TB, NB>(msp, tid, op_begin, op_count, oc, b);
        Cx<R> v[N];
#pragma unroll
        for (int s = 0; s < N; s++) v[s] = lds_at(tile, b ^ reg_offset<NB>(s, oc));
        const uint32_t f = reg_step_ops<R, N, LIGHT>(v, msp, op_begin, op_count, ops, coef);
        // an out-of-place pass writes the other buffer, at its own tile base
        Cx<R>* __restrict__ out = pa.scatter ? dst : state;
        const uint64_t ob = pa.scatter ? out_base(pa, blk, sa + pa.scatter_idx) : base;
        const uint64_t gs = ob ^ direct_thread_offset<TB, NB>(pa.last_g, tid) ^
                            direct_reg_offset<NB>(pa.last_g, (int)f);
#pragma unroll
        for (int s = 0; s < N; s++) out[gs ^ direct_reg_offset<NB>(pa.last_g, s)] = v[s];
    } else if (pa.scatter) {
        scatter_store<R, TB, NB>(dst, pa, blk, tile, sa);
    } else {
        store_tile<R, TB, NB>(state, pa, base, tile);
    }
}

// Waves per SIMD that the LDS tiles of one CU (160 KiB) allow.
template <typename R, int TB, int NB>
constexpr int kLdsWavesPerSimd =
    (163840 / ((1 << TB) * 2 * (int)sizeof(R))) * ((1 << (TB - NB)) / 64) / 4;

// Minimum waves per SIMD the register allocation must allow: two for
// complex128 16-amplitude threads (their 64 KiB LDS tiles allow two per CU),
// for the other light variants as many as their LDS tiles allow, at most
// four; the compiler's own choice otherwise.
template <typename R, int TB, int NB, bool LIGHT>
constexpr int kPassWavesPerEU =
    (sizeof(R) == 8 && NB == 4) ? 2
                                : (LIGHT ? (kLdsWavesPerSimd<R, TB, NB> < 1 ? 1
                                            : (kLdsWavesPerSimd<R, TB, NB> > 4 ? 4 : kLdsWavesPerSimd<R, TB, NB>))
                                         : 1);

// LIGHT: the pass's ops are all in the light set (apply_op_light).
template <typename R, int TB, bool WIDE, int NB, bool LIGHT>
__global__ __launch_bounds__(1 << (TB - NB)) __attribute__((amdgpu_waves_per_eu(kPassWavesPerEU<R, TB, NB, LIGHT>)))
void k_pass(Cx<R>* __restrict__ state, PassArgs pa,
                                                         const MicroStep* __restrict__ micro,
                                                         const MicroOp* __restrict__ ops,
                                                         const Cx<R>* __restrict__ coef,
                                                         const uint32_t* __restrict__ raw,
                                                         Cx<R>* __restrict__ dst,
                                                         const ScatterArgs* __restrict__ sa)
{
    __shared__ Cx<R> tile[1 << TB];
    tile_pass<R, TB, WIDE, NB, LIGHT>(state, pa, blockIdx.x, tile, micro, ops, coef, raw, dst, sa);
}

// ---- many circuits in one persistent launch ---------------------------------
//
// Tasks are (circuit, pass, tile), queued by pass level, then circuit, then
// tile.  A workgroup takes the next task; a task of pass p > 0 first waits
// until every tile of its circuit's pass p-1 is done.  Tasks are taken in
// queue order and only wait on earlier ones, which running workgroups hold,
// so the queue always drains; the wait is still bounded, and a wait that
// runs out sets *err instead of spinning on.
// complex64 register-step batch kernels are held to 128 VGPRs (4 waves per
// SIMD): with 137 the 8-wave workgroups of 13-bit tiles ran one per CU.
template <typename R, bool WIDE>
constexpr int kBatchWavesPerEU = (sizeof(R) == 4 && !WIDE) ? 4 : 1;

template <typename R, int TB, bool WIDE, int NB>
__global__ __launch_bounds__(1 << (TB - NB)) __attribute__((amdgpu_waves_per_eu(kBatchWavesPerEU<R, WIDE>)))
void k_batch(Cx<R>* __restrict__ states,
                                                         const uint64_t* __restrict__ circ_off,
                                                         const BatchTask* __restrict__ tasks, uint32_t ntasks,
                                                         uint32_t* __restrict__ next, uint32_t* __restrict__ done,
                                                         const PassArgs* __restrict__ passes,
                                                         const MicroStep* __restrict__ micro,
                                                         const MicroOp* __restrict__ ops,
                                                         const Cx<R>* __restrict__ coef,
                                                         const uint32_t* __restrict__ raw, uint32_t* __restrict__ err)
{
    __shared__ Cx<R> tile[1 << TB];
    __shared__ uint32_t s_task;
    // One thread-0 block per iteration, between two barriers: it counts the
    // task just finished and fetches the next one.  Two separate thread-0
    // blocks (count at the end of an iteration, fetch at the start of the
    // next) sit side by side across the back edge with no barrier between
    // them; the compiler merges them into a thread-0-only path around the
    // loop head and leaves the other lanes of wave 0 circling an inner loop
    // that re-reads the old task index forever.  The index is also read
    // with readfirstlane so that the exit test is a scalar branch.
    //
    // Cross-workgroup visibility costs one agent-scope release (before a
    // finished task is counted) and one agent-scope acquire (after a wait),
    // both by thread 0: on this GPU they write back / invalidate the whole
    // L2 of the XCD (and the CU's L1), which every wave of the workgroup
    // shares, so one wave's fence covers the workgroup once the barrier
    // after the tile has drained every wave's stores.  (Fencing in every
    // thread made the batch four times slower.)
    uint32_t finished = 0xFFFFFFFFu;
    for (;;) {
        if (threadIdx.x == 0) {
            if (finished != 0xFFFFFFFFu) {
                __threadfence();
                __builtin_amdgcn_s_waitcnt(0);  // the L2 write-back has finished
                atomicAdd(done + finished, 1u);
            }
            s_task = atomicAdd(next, 1u);
        }
        __syncthreads();
        const uint32_t t = __builtin_amdgcn_readfirstlane(s_task);
        if (t >= ntasks) break;
        const BatchTask task = tasks[t];
        if (task.wait_pass != 0xFFFFFFFFu) {
            if (threadIdx.x == 0) {
                uint32_t spins = 0;
                while (__hip_atomic_load(done + task.wait_pass, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                       task.wait_count) {
                    __builtin_amdgcn_s_sleep(4);
                    if (++spins > (1u << 20)) {
                        atomicOr(err, 1u);
                        break;
                    }
                }
                __threadfence();
                __builtin_amdgcn_s_waitcnt(0);  // the cache invalidation has finished
            }
            __syncthreads();
        }
        tile_pass<R, TB, WIDE, NB, false>(states + circ_off[task.circuit], passes[task.pass], task.tile, tile, micro, ops,
                                          coef, raw, nullptr, nullptr);
        // every wave's stores of this tile have completed (each wave drains
        // its memory counters; a workgroup-scope barrier alone does not wait
        // for global stores) before thread 0 fences and counts the task, and
        // every wave has read s_task before thread 0 overwrites it
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        // the last pass of a circuit is read only after this launch ends:
        // no task waits on it, so it is neither fenced nor counted
        finished = task.signal ? task.pass : 0xFFFFFFFFu;
    }
}

template <typename R, int TB, int NB>
int pass_t(void* state, int nq_int, const PassArgs& pa, const DevPlan& dp, hipStream_t s)
{
    // one workgroup per tile; a run from |0...0> may list fewer rest qubits
    (void)nq_int;
    const unsigned grid = 1u << pa.nrest;
    if (pa.has_wide)
        hipLaunchKernelGGL((k_pass<R, TB, true, NB, false>), dim3(grid), dim3(1 << (TB - NB)), 0, s, (Cx<R>*)state,
                           pa, dp.micro, dp.ops, (const Cx<R>*)dp.coef, dp.raw, (Cx<R>*)dp.dst, dp.scatter);
    else if (pa.light)
        hipLaunchKernelGGL((k_pass<R, TB, false, NB, true>), dim3(grid), dim3(1 << (TB - NB)), 0, s, (Cx<R>*)state,
                           pa, dp.micro, dp.ops, (const Cx<R>*)dp.coef, dp.raw, (Cx<R>*)dp.dst, dp.scatter);
    else
        hipLaunchKernelGGL((k_pass<R, TB, false, NB, false>), dim3(grid), dim3(1 << (TB - NB)), 0, s, (Cx<R>*)state,
                           pa, dp.micro, dp.ops, (const Cx<R>*)dp.coef, dp.raw, (Cx<R>*)dp.dst, dp.scatter);
    return (int)hipGetLastError();
}

// complex128 runs 3- or 4-bit micro-steps, complex64 4-bit ones.
template <typename R, int TB>
int pass_nb(void* state, int nq_int, const PassArgs& pa, const DevPlan& dp, hipStream_t s)
{
    if constexpr (sizeof(R) == 8)
        if (pa.micro_bits == 3) return pass_t<R, TB, 3>(state, nq_int, pa, dp, s);
    if (pa.micro_bits != 4) return (int)hipErrorInvalidValue;
    return pass_t<R, TB, 4>(state, nq_int, pa, dp, s);
}

template <typename R>
int pass_r(int tile_bits, void* state, int nq_int, const PassArgs& pa, const DevPlan& dp, hipStream_t s)
{
    switch (tile_bits) {
    case 10: return pass_nb<R, 10>(state, nq_int, pa, dp, s);
    case 11: return pass_nb<R, 11>(state, nq_int, pa, dp, s);
    case 12: return pass_nb<R, 12>(state, nq_int, pa, dp, s);
    case 13: return pass_nb<R, 13>(state, nq_int, pa, dp, s);
    default: return (int)hipErrorInvalidValue;
    }
}

template <typename R, int TB, bool WIDE, int NB>
int batch_t(const BatchArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL((k_batch<R, TB, WIDE, NB>), dim3(a.grid), dim3(1 << (TB - NB)), 0, s, (Cx<R>*)a.states,
                       a.circ_off, a.tasks, a.ntasks, a.next, a.done, a.passes, a.dp.micro, a.dp.ops,
                       (const Cx<R>*)a.dp.coef, a.dp.raw, a.err);
    return (int)hipGetLastError();
}

template <typename R, int TB, bool WIDE>
int batch_nb(const BatchArgs& a, hipStream_t s)
{
    if constexpr (sizeof(R) == 8)
        if (a.micro_bits == 3) return batch_t<R, TB, WIDE, 3>(a, s);
    if (a.micro_bits != 4) return (int)hipErrorInvalidValue;
    return batch_t<R, TB, WIDE, 4>(a, s);
}

template <typename R, bool WIDE>
int batch_w(const BatchArgs& a, hipStream_t s)
{
    switch (a.tile_bits) {
    case 10: return batch_nb<R, 10, WIDE>(a, s);
    case 11: return batch_nb<R, 11, WIDE>(a, s);
    case 12: return batch_nb<R, 12, WIDE>(a, s);
    case 13: return batch_nb<R, 13, WIDE>(a, s);
    default: return (int)hipErrorInvalidValue;
    }
}

}  // namespace
}  // namespace kate
