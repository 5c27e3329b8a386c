// plan.hpp -- host-side program and pass plan shared by the planner and the
// launchers.  Plain data only (no HIP types), so the planner compiles as
// ordinary C++ and can be unit-tested on the CPU.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace kate {

// Register micro-step op kinds (op code = kind << 8 | variant).
enum OpKind : uint16_t {
    OK_DENSE = 0,  // variant = 4-bit mask of micro bits (ascending = matrix bit order)
    OK_PHASE = 1,  // diagonal 1-qubit gate, variant = micro bit
    OK_H = 2,      // variant = micro bit
    OK_CNOT = 3,   // variant = target*4 + control
    OK_CRK = 4,    // variant = a*4 + b, a < b
    OK_TOFF = 5,   // variant = target*16 + c0*4 + c1, c0 < c1
    OK_KRON = 6,   // variant = mask of micro bits; one 2x2 per set bit, ascending, applied in turn
    OK_EKRON = 7,  // variant = mask of micro bits (2+ bits) | 16 if the right diagonal is present;
                   // the same product with every 2x2 in Euler form D1 * RX * D2
    OK_FKRON = 8,  // variant = mask of micro bits | flip mask << 4: OK_EKRON without D2 on every
                   // register bit, each RX scaled to [[1, -it], [-it, 1]] (|t| <= 1; scales in the
                   // D1 table); the step's store writes register s to the place of s ^ flip
};

enum StepKind : uint8_t {
    SK_REG = 0,      // 4 tile bits held in registers, ops applied in place
    SK_GENERIC = 1,  // one dense gate of any size, matrix-vector on the LDS tile
    SK_MFMA = 2,     // one dense gate of 4 or 5 qubits on matrix cores (complex64, FP32 products)
    SK_MFMA_BF16 = 3,  // the same with split-bf16 products (x = hi + lo, three products), FP32 accumulation
};

struct MicroOp {
    uint16_t code;
    uint16_t pad;
    uint32_t coef;  // complex-element offset into the coefficient pool
};

struct MicroStep {
    uint8_t kind;
    uint8_t nbits;
    uint16_t op0_code;  // SK_REG: code of ops[op_begin], read with the step (no dependent load)
    uint32_t op_begin;
    uint32_t op_count;
    uint32_t coef;     // SK_GENERIC: matrix offset (complex elements); SK_MFMA*: operand offset in
                       // Plan::raw; SK_REG: coefficient offset of ops[op_begin]
    uint8_t bits[16];  // SK_REG: ascending tile bits; SK_GENERIC/SK_MFMA: tile bit of matrix bit i
    // Layout of the LDS tile when the step runs: the amplitude of tile index
    // y sits at LDS index XOR_{b in y} cols[b] (linear over GF(2); CNOTs are
    // applied by updating cols instead of moving data).
    uint16_t cols[16];
    // The same layout pre-composed for SK_REG addressing: thread-index bit i
    // (the i-th tile bit outside bits[0..4)) contributes tcols[i], register
    // bit k contributes pcols[k].
    uint16_t tcols[12];
    uint16_t pcols[4];
    // SK_MFMA: layout columns of matrix bit i (gcols) and, in tcols, of the
    // non-gate tile bits in ascending order (the columns of the contraction).
    uint16_t gcols[8];
    // SK_REG: tcols and pcols as LDS byte offsets after the tile's XOR
    // swizzle (the swizzle is linear over GF(2), so the byte offset of a
    // register amplitude is the XOR of these).
    uint32_t tbyte[12];
    uint32_t pbyte[4];
};

constexpr int kMaxQubits = 40;

// One pass over the state: tile bit b of the LDS tile is state qubit
// load_pos[b] when the tile is read and store_pos[b] when it is written back
// (they differ only for qubit remaps).  rest_pos lists the other qubits in
// ascending order; block index bit i is qubit rest_pos[i].
struct PassArgs {
    uint8_t tile_bits;
    uint8_t nrest;
    uint8_t load_pos[16];
    uint8_t store_pos[16];
    uint8_t rest_pos[kMaxQubits];
    uint8_t has_wide;    // some micro-step is not SK_REG (selects the kernel variant)
    uint8_t micro_bits;  // bits per register micro-step (3 or 4): 2^micro_bits amplitudes per thread
    uint32_t micro_begin;
    uint32_t micro_count;
    uint16_t store_cols[16];  // tile layout at the end of the pass (see MicroStep::cols)
    uint32_t store_byte[16];  // the same as swizzled LDS byte offsets
    // Direct first / last micro-step: the first SK_REG step loads its
    // amplitudes from HBM into registers (no LDS fill), the last one stores
    // them to HBM from registers (no LDS drain).  first_g[i] / last_g[i] is
    // the state offset contributed by thread-index bit i of that step
    // (i < tile_bits - micro_bits), first_g[12 + k] / last_g[12 + k] that of
    // register bit k; offsets combine by XOR with the tile base.
    uint8_t fuse_first;
    uint8_t fuse_last;
    // Set per launch on the first pass of an execute from a basis state: the
    // tile is not read from HBM but made in registers (amplitude 1 at state
    // offset synth_index, 0 elsewhere), which replaces the init pass.
    uint8_t synth_basis;
    // Every op of the pass is in the light set (Euler-form Kronecker ops
    // without a right diagonal, dense 1-qubit ops, phases, Hadamards,
    // CNOTs): the pass runs on the kernel variant that dispatches only those.
    uint8_t light;
    // An out-of-place pass: the tile is read from one state buffer and its
    // amplitudes written to 2^low-amplitude runs of the other, through
    // Plan::scatter_args[scatter_idx].
    uint8_t scatter;
    uint64_t first_g[16];
    uint64_t last_g[16];
    uint64_t synth_index;
    // Set per launch in a run from |0...0>: amplitudes whose state offset
    // meets zero_mask are known to be zero and are not read.
    uint64_t zero_mask;
    uint32_t scatter_idx;
    uint32_t scatter_pad;
};

// Store side of a scatter pass: store index bit k of a tile goes to state
// offset out_tile[k], bit i of the tile number to out_rest[i] (offsets
// combine by XOR); store index bits 0..low-1 are offset bits 0..low-1, so
// every write is a whole run.
struct ScatterArgs {
    uint64_t out_tile[16];
    uint64_t out_rest[kMaxQubits];
};

struct PlanGate {
    int id = 0;
    int k = 0;
    int q[16] = {0};
    int optarg = 0;
    int64_t moff = -1;  // complex offset into Program::mat
    uint64_t mask = 0;
};

// Validated copy of a kate_program.
struct Program {
    int nqubits = 0;
    std::vector<PlanGate> gates;
    std::vector<double> mat;  // interleaved complex128
    std::vector<int> group_end;
    std::vector<uint64_t> group_mask;
    // instruction stream: kind 1 = compute next group, kind 0 = swap pairs
    std::vector<int> instr_kind;
    std::vector<std::vector<int>> instr_pairs;
    std::vector<int> initial_permutation;
    std::vector<int> final_inverse_permutation;
};

struct PlanOptions {
    int tile_bits = 12;   // qubits per LDS tile
    int low_qubits = 3;   // qubits 0..low-1 are in every gate tile
    int min_qubits = 12;  // the state is padded to at least this many qubits
    bool mfma = false;       // SK_MFMA for dense gates of 4..5 qubits (complex64 only)
    bool mfma5 = false;      // SK_MFMA for dense gates of 5 qubits only (complex64 only)
    bool mfma_fuse = false;  // also compose the gates of 5-qubit windows into SK_MFMA steps
    bool mfma_bf16 = false;  // matrix-core steps use split-bf16 products (SK_MFMA_BF16)
    int micro_bits = 4;      // register micro-steps span 3 or 4 tile bits (8 or 16 amplitudes per thread)
    int amp_bytes = 16;      // 16 (complex128) or 8 (complex64): LDS byte offsets in the plan
    bool euler = true;       // Kronecker ops of 2+ composed 1-qubit gates in Euler form (OK_EKRON)
    bool factor = true;      // a step's lone Euler-form op on every register bit scaled to OK_FKRON
    bool direct = true;      // first / last SK_REG step of a pass loads / stores HBM directly
    bool auto_low = false;   // build the plan for several low_qubits and keep the cheapest
    bool auto_tile = false;  // complex64: 13-qubit tiles for all-light plans, else 12 (tile_bits ignored)
    bool support = true;     // runs from |0...0> skip tiles known to be zero (used by the simulator)
    bool defer = true;       // execute() measures through the trailing permutation passes (simulator)
    bool split = true;       // dense 2-qubit gates that are CNOT x (A (x) B) products run as their factors
    bool reorder = true;     // swap-free programs: also plan the gates in a canonical order, keep the cheaper
    bool scatter = false;    // 2+ trailing permutation passes become one out-of-place pass (needs a second buffer)
    bool relabel = false;    // with scatter: passes after the support region read contiguous tiles (relabelled qubits)
    bool zero_start = true;  // the first pass sees |0...0> (false: any state, e.g. a product input; no support region)
};

struct Plan {
    int nq_int = 0;  // padded qubit count the kernels see
    int tile_bits = 0;
    std::vector<PassArgs> passes;
    std::vector<uint8_t> pass_is_remap;
    std::vector<MicroStep> micro;
    std::vector<MicroOp> ops;
    std::vector<double> coef;  // interleaved complex128; converted per dtype on upload
    std::vector<uint32_t> raw; // matrix-core operands, uploaded bit for bit
    int gate_passes = 0;
    int remap_passes = 0;
    int low_qubits = 0;  // qubits 0..low-1 in every gate tile (as planned)
    // Support of a run that starts from |0...0>: pass_zero[i] holds the
    // qubits q for which every amplitude with bit q set is exactly zero when
    // pass i starts (no gate has yet moved weight onto q = 1).  Passes
    // before the first remap pass (support_passes of them) may skip the
    // tiles whose non-tile qubits meet that set.
    std::vector<uint64_t> pass_zero;
    int support_passes = 0;
    // Trailing passes [trail_begin, passes.size()) only permute amplitudes
    // (lazy-CNOT layouts or remaps, no micro-steps): together a GF(2)-linear
    // map A of state offsets.  trail_inv[q] = A^-1(2^q), the offset before
    // those passes of the amplitude that ends at offset 2^q; a measurement
    // can read the state through it instead of running them.
    size_t trail_begin = 0;
    std::vector<uint64_t> trail_inv;
    // The pass that applies the trailing permutation out of place (-1 if
    // none), and the store offsets of every out-of-place pass.
    long scatter_pass = -1;
    std::vector<ScatterArgs> scatter_args;
    // First pass of the relabelled region: from there every pass loads
    // its tile from state bits 0..tile_bits-1 and stores, out of place, to
    // the positions the next pass reads (the last to virtual order).
    long relabel_begin = -1;
    std::string describe(bool verbose = false) const;
};

// Builds the pass plan.  Returns an empty string on success, else the error.
std::string build_plan(const Program& prog, const PlanOptions& opt, Plan& out);

// Launch arguments of pass i of a run that starts from |0...0> without an
// init kernel: pass 0 makes its tile in registers; with `support`, a pass
// reads no amplitude known to be zero (Plan::pass_zero) and, before the
// first remap pass and unless it is the last pass, launches only the tiles
// whose rest qubits are all outside that set (fewer rest qubits).  n_run is
// the number of passes the run launches (the last of them writes every tile).
PassArgs zero_start_pass(const Plan& plan, size_t i, bool support, size_t n_run);

}  // namespace kate
