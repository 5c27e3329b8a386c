// kernels.hpp -- launchers of the gfx950 kernels (kernels.hip).  Every
// launcher enqueues on `stream` and returns the hipError_t of the launch as
// an int (0 = success).
#pragma once

#include <cstdint>

#include "plan.hpp"

namespace kate {

struct DevPlan {
    const MicroStep* micro;
    const MicroOp* ops;
    const void* coef;  // complex of the state's precision
    const uint32_t* raw;  // matrix-core operands
    void* dst = nullptr;                  // a gate pass with PassArgs::scatter writes this buffer
    const ScatterArgs* scatter = nullptr; // through these offsets (device memory)
};

struct BatchTask {
    uint32_t circuit;
    uint32_t pass;        // global pass index
    uint32_t tile;
    uint32_t wait_pass;   // pass that must be complete first (0xFFFFFFFF: none)
    uint32_t wait_count;  // its tile count
    uint32_t signal;      // 1: a later pass of the circuit waits on this one (count it, after a release)
    uint32_t pad[2];
};

struct BatchArgs {
    void* states;
    uint64_t total_amps;
    const uint64_t* circ_off;  // amplitude offset of each circuit's state
    const int32_t* nq;
    const int32_t* nq_int;
    const int64_t* zoff;
    double* z;
    double* norms;
    uint32_t count;
    const BatchTask* tasks;
    uint32_t ntasks;
    uint32_t* next;
    uint32_t* done;
    uint32_t npasses;
    uint32_t* err;
    const PassArgs* passes;
    DevPlan dp;
    int tile_bits;
    int micro_bits;  // amplitudes per thread = 2^micro_bits
    bool wide;
    unsigned grid;
};

// Init every state, run every task of the batch in one persistent launch,
// then <Z_q> and norms per circuit.  All kernels, no memsets: the sequence
// can be captured in a graph.
int launch_batch(int dtype, const BatchArgs& a, void* stream);

// One pass (gate application or remap) over the state.
int launch_pass(int dtype, int tile_bits, void* state, int nq_int, const PassArgs& pa, const DevPlan& dp,
                void* stream);
// Calibration: every tile of the pass read and written back in place
// (load positions, no arithmetic, no LDS): the HBM time of its access pattern.
int launch_copy_tiles(int dtype, int tile_bits, void* state, int nq_int, const PassArgs& pa, void* stream);
// The out-of-place permutation pass (Plan::scatter_pass): every tile of `src`
// to runs of `dst` through `sa` (device memory); for a gate pass with
// PassArgs::scatter, a plain copy through its store (the copy ceiling).
int launch_scatter(int dtype, int tile_bits, const void* src, void* dst, const PassArgs& pa, const ScatterArgs* sa,
                   void* stream);
// |index> (basis state).
int launch_init_basis(int dtype, void* state, int nq_int, uint64_t index, void* stream);
// Product state: factor[b*nq + j] (complex128) for bit value b of qubit j.
int launch_init_product(int dtype, void* state, int nq_int, const double* d_factors, void* stream);
// Multiplies every amplitude by `factor`.
int launch_scale(int dtype, void* state, int nq_int, double factor, void* stream);

// Sizes of the scratch buffer (in doubles) the reductions need.
uint64_t reduce_scratch_doubles(int nq_int);
// Amplitude sums: out[0..1] = sum of all amplitudes, out[2+2q..] = sum of the
// amplitudes whose bit q is 1, for q < nq_int (complex128, FP64 accumulation,
// fixed summation order).
// gather_cols (nullable, nq_int entries): read index i at the GF(2)-linear
// offset XOR_{q in i} gather_cols[q] (a state whose trailing permutation
// passes have not run).
int launch_amp_sums(int dtype, const void* state, int nq_int, double* d_scratch, double* d_out, void* stream,
                    const uint64_t* gather_cols = nullptr);
// out[j] = sum_x |a_x|^2 (-1)^popcount(x & masks[j]) for j < count (FP64).
int launch_zsums(int dtype, const void* state, int nq_int, const uint64_t* d_masks, int count, double* d_scratch,
                 double* d_out, void* stream);

// d_out[0..3] = {Re, Im} of <a|b> = sum conj(a_i) b_i, sum |a_i|^2,
// sum |b_i|^2 (FP64, fixed order); d_out[4] = max |a_i - b_i|.  The two
// states may differ in precision.
int launch_overlap(int dtype_a, const void* a, int dtype_b, const void* b, int nq_int, double* d_scratch,
                   double* d_out, void* stream);

}  // namespace kate
