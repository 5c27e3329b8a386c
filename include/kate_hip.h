/*
 * kate_hip.h -- C ABI of the MI355X state-vector library (libkate_hip.so).
 *
 * Kate (the reference) exposes only a header-only C++ API; it has no C ABI,
 * plugin registry or FFI.  Every entry point below replaces one step of the
 * reference's single-GPU simulate call, cited per declaration.  The C++
 * drop-in layer (kate_amd/kate/Kate.hpp) lowers Kate's QuantumCircuit onto
 * these calls; ctypes (kate_amd/__init__.py) binds the same symbols.
 *
 * Conventions: plain pointers and sizes, int status (KATE_OK = 0, negative on
 * error, message in kate_last_error()), no exceptions across the boundary.
 * Amplitudes are interleaved complex (re, im): complex128 (16 B) or complex64
 * (8 B).  Bit q of an amplitude index is the value of qubit q.
 */
#ifndef KATE_HIP_H
#define KATE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KATE_MAX_QUBITS 40
#define KATE_MAX_GATE_QUBITS 16

enum kate_status {
    KATE_OK = 0,
    KATE_ERR_ARG = -1,
    KATE_ERR_NO_DEVICE = -2,
    KATE_ERR_HIP = -3,
    KATE_ERR_OOM = -4,
    KATE_ERR_PROGRAM = -5,
    KATE_ERR_UNSUPPORTED = -6,
    KATE_ERR_BUSY = -7
};

enum kate_dtype { KATE_C128 = 0, KATE_C64 = 1 };

/* Gate ids: backendagnosticcode/agnosticpreprocessor.hpp:28-33. */
enum kate_gate_id { KATE_DENSE = 0, KATE_H = 2, KATE_CNOT = 3, KATE_CRK = 4, KATE_TOFFOLI = 5 };

/* One gate of the flat program: the lowering of Kate::Gate
 * (backendagnosticcode/basic_host_types.hpp:190-296).  qubits[i] is the
 * qubit of matrix-index bit i (order is semantic); matrix_offset counts
 * complex elements into kate_program.matrix_pool (-1 when not dense). */
typedef struct kate_gate_rec {
    int32_t id;
    int32_t nq;
    int32_t qubits[KATE_MAX_GATE_QUBITS];
    int32_t optarg;
    int32_t reserved;
    int64_t matrix_offset;
} kate_gate_rec;

/* The flat lowering of Kate::QuantumCircuit (QuantumCircuit.hpp:12-17):
 * gates in execution order; groups[g] = (end-exclusive gate index, qubit
 * bitmask); instruction i is a compute (kind 1, "run the next group") or a
 * swap list (kind 0) whose position pairs are
 * pairs[instr_pair_offset[i] .. + 2*instr_npairs[i]); permutations as Kate
 * stores them (initial: real -> virtual, final inverse: virtual -> real).
 * NULL permutations mean identity; ninstr == 0 means "one compute per group".
 * The library copies everything it needs; the caller keeps ownership. */
typedef struct kate_program {
    int32_t nqubits;
    int32_t ngates;
    const kate_gate_rec* gates;
    const double* matrix_pool; /* interleaved complex128 */
    int64_t matrix_pool_len;   /* complex elements */
    int32_t ngroups;
    const int32_t* group_end;
    const uint64_t* group_mask;
    int32_t ninstr;
    const int32_t* instr_kind;
    const int32_t* instr_npairs;
    const int32_t* instr_pair_offset;
    const int32_t* pairs;
    const int32_t* initial_permutation;
    const int32_t* final_inverse_permutation;
} kate_program;

/* Execution options (no counterpart in Kate, whose kernels have one shape). */
typedef struct kate_options {
    int32_t dtype;        /* kate_dtype; KATE_C128 is the plain path */
    int32_t tile_qubits;  /* 10..13; 0 = 12 for KATE_C128; for KATE_C64 the
                             planner takes 13 or 12 per circuit */
    int32_t low_qubits;   /* qubits always in a tile (coalesced run = 2^low amplitudes); 0 = the
                             planner tries 2-5 (complex128) / 3-6 (complex64) and keeps the plan
                             with the lowest estimated HBM + register-step cost */
    int32_t use_mfma;     /* C64 only: 1 = dense gates of 4-5 qubits on the matrix cores;
                             2 = also compose the gates of each 5-qubit window into one
                             dense gate (FP64 composition) for the matrix cores;
                             3 = as 2 with split-bf16 products (x = hi + lo, three
                             products per term) and FP32 accumulation */
    int32_t batch_wgs_per_cu; /* kate_batch_*: workgroups per CU of the persistent launch; 0 = 8 */
    int32_t device;       /* HIP device ordinal */
    int32_t micro_bits;   /* amplitudes per thread in register steps = 2^micro_bits:
                             3 (C128 only) or 4; 0 = default per dtype */
    int32_t plan_flags;   /* 0 = every plan optimisation; KATE_PLAN_* bits switch one off */
} kate_options;

/* kate_options.plan_flags */
enum {
    /* apply composed 1-qubit gates as dense 2x2 products instead of writing
       products of 2-4 of them as D1 * (RX x ... x RX) * D2 (Euler form:
       fewer FMAs, same result to rounding) */
    KATE_PLAN_NO_EULER = 1,
    /* fill and drain every LDS tile through LDS instead of loading the
       first micro-step's amplitudes from HBM into registers and storing the
       last one's from registers (same arithmetic, bit for bit) */
    KATE_PLAN_NO_DIRECT = 2,
    /* execute() from |0...0>: read and launch every tile instead of only
       those that can hold non-zero amplitudes (qubits no gate has touched
       yet keep every amplitude with their bit set at exactly zero) */
    KATE_PLAN_NO_SUPPORT = 4,
    /* execute(): run the trailing passes that only permute amplitudes
       (CNOT layers at the end of a circuit) instead of measuring through
       them and running them when the state is next read */
    KATE_PLAN_NO_DEFER = 8,
    /* apply dense 2-qubit gates that factor as (CNOT x) (A (x) B) (x CNOT)
       (what Kate's gateFusion makes of two neighbouring 1-qubit gates and
       the CNOT between them) as dense 4x4 products instead of as their
       factors (1-qubit gates composed with their neighbours, the CNOT a
       layout update; same result to rounding) */
    KATE_PLAN_NO_SPLIT = 16,
    /* plan the gates of a swap-free program only in the order given, instead
       of also in a canonical order of the same dependency graph (ASAP level,
       then lowest qubit) and keeping the cheaper plan */
    KATE_PLAN_NO_REORDER = 32,
    /* keep Euler-form Kronecker steps in the D1 * RX form instead of
       scaling each rotation by its larger entry (one FMA per output component
       instead of a multiply and an FMA; the scales go into the diagonal
       table, and a rotation whose off-diagonal entry is the larger one is
       applied with its two outputs exchanged, which the step's store undoes;
       same result to rounding) */
    KATE_PLAN_NO_FACTOR = 64,
    /* run two or more trailing permutation passes (the CNOT layers at the
       end of a circuit) as they are planned instead of as one out-of-place
       pass into a second state buffer (which is allocated only when the
       GPU has room for both states and 1 GiB more; results bit for bit the
       same) */
    KATE_PLAN_NO_SCATTER = 128,
    /* keep every pass in place on the virtual qubit order instead of, from
       the first pass after which no qubit is known zero, reading each tile
       from the lowest state bits and writing (out of place, into the second
       buffer) where the next pass reads (the last pass to virtual order;
       plans of 12-qubit tiles or narrower; needs the second buffer as
       KATE_PLAN_NO_SCATTER describes; results bit for bit the same) */
    KATE_PLAN_NO_RELABEL = 256,
    /* keep dense 5-qubit gates of a KATE_C64 plan on the VALU (register
       blocks) instead of the FP32 matrix cores, which take them by default
       (use_mfma = 0); results to FP32 rounding */
    KATE_PLAN_NO_MATRIX5 = 512
};

/* Wall-clock split of the last execute, as Kate's execute(true) prints it
 * (HIPcode/simulator.hpp:301-353), plus the pass statistics of the plan. */
typedef struct kate_timings {
    double init_ms;
    double compute_ms;
    double measure_ms;
    int32_t passes;        /* gate-application passes over the state */
    int32_t remap_passes;  /* qubit-remap passes (Kate's swap instructions) */
    int32_t micro_steps;   /* register micro-steps inside tiles, summed over passes */
    int32_t gates;         /* gates in the program */
    double bytes_per_pass; /* 2 * 2^n * sizeof(amplitude) */
} kate_timings;

typedef struct kate_sim kate_sim;
typedef struct kate_circuit kate_circuit;

/* ---- library ---- */
const char* kate_status_string(int status);
/* Thread-local message of the last failing call on this thread. */
const char* kate_last_error(void);
/* HIPcode/simulator.hpp:244-259 (hipGetDeviceCount).  No CPU fallback. */
int kate_device_count(int* count);
/* HIPcode/DeviceInfo.hpp:10-36 (printGpuInfo): one line per device into buf. */
int kate_device_info(char* buf, size_t len);

/* ---- simulator: HIPcode/simulator.hpp:220-543,760-810 ---- */
/* Simulator::Simulator (HIPcode/simulator.hpp:243-300): validates and copies
 * the program, plans its passes, allocates the state (hipMalloc of 2^n
 * amplitudes) and uploads the whole program once. */
int kate_sim_create(const kate_program* program, const kate_options* options, kate_sim** out);
/* Simulator::~Simulator (HIPcode/simulator.hpp:407-425). */
int kate_sim_destroy(kate_sim* sim);
/* initialize() / initialize_state (HIPcode/simulator.hpp:427-443,17-24),
 * generalised to any basis index (virtual order). */
int kate_sim_init_basis(kate_sim* sim, uint64_t index);
/* initialize(proba_state&) / initialize_probastate (HIPcode/simulator.hpp:444-482,26-38):
 * theta_phi[2*i], theta_phi[2*i+1] for real qubit i. */
int kate_sim_init_product(kate_sim* sim, const double* theta_phi);
/* The instruction loop of execute() (HIPcode/simulator.hpp:321-339): every
 * compute group and swap instruction, enqueued on the simulator's stream. */
int kate_sim_run(kate_sim* sim);
/* measurement() (HIPcode/simulator.hpp:483-543): Kate's proba_state,
 * theta_phi[2*q], theta_phi[2*q+1] for real qubit q. */
int kate_sim_measure_proba(kate_sim* sim, double* theta_phi);
/* execute(displaytime) / execute(proba_state&, displaytime)
 * (HIPcode/simulator.hpp:301-406): init (basis |0...0> when theta_phi_in is
 * NULL, else product state) + run + measurement, timed like Kate's. */
int kate_sim_execute(kate_sim* sim, const double* theta_phi_in, double* theta_phi_out);
/* No counterpart in Kate: <Z_S> = sum_x |a_x|^2 (-1)^popcount(x & S) for each
 * real-qubit mask S, accumulated in FP64, normalised by the state's norm. */
int kate_sim_expect_z(kate_sim* sim, const uint64_t* masks, int32_t count, double* out);
/* No counterpart in Kate, and no allocating counterpart either: the
 * kate_sim_expect_z reduction for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  d_state: a caller-owned device state of 2^nqubits amplitudes
 * (dtype, virtual order, 12 <= nqubits <= 40) on `device`; masks: `count`
 * (1..15) Z strings over state bits (host array); d_out: `count` doubles of
 * device memory receiving <Z_S> = sum_x |a_x|^2 (-1)^popcount(x & S) / norm,
 * FP64.  Enqueued on `stream` (asynchronous).  The partial sums live in a
 * fixed 256 KiB device array that the library holds on each GPU until the
 * process ends; calls that overlap on one GPU share it and race. */
int kate_state_expect_z_noalloc(int32_t device, int32_t dtype, const void* d_state, int32_t nqubits,
                                const uint64_t* masks, int32_t count, double* d_out, void* stream);
/* No counterpart in Kate: sum_x |a_x|^2 in FP64. */
int kate_sim_norm(kate_sim* sim, double* out);
/* No counterpart in Kate: rescale the state to norm 1 (factor in FP64). */
int kate_sim_renormalize(kate_sim* sim);
/* Public member Simulator::gpu_qbits_states (HIPcode/simulator.hpp:232): the
 * amplitudes, copied to host memory, in virtual order (real_order = 0, as
 * Kate leaves them) or real order (real_order = 1, final_inverse_permutation
 * applied).  bytes must be 2^n * sizeof(amplitude). */
int kate_sim_read_state(kate_sim* sim, void* host, uint64_t bytes, int32_t real_order);
/* No counterpart in Kate: out[0..1] = <a|b> = sum conj(a_x) b_x, out[2] =
 * <a|a>, out[3] = <b|b>, all FP64 (fidelity = |<a|b>|^2 / (<a|a><b|b>)),
 * out[4] = max_x |a_x - b_x|.  The two simulators may differ in precision;
 * sizes and qubit orders must match. */
int kate_sim_overlap(kate_sim* a, kate_sim* b, double* out5);
/* No counterpart in Kate: overwrite the state (virtual order). */
int kate_sim_write_state(kate_sim* sim, const void* host, uint64_t bytes);
/* Device address of the state (Simulator::gpu_qbits_states[0]).  A plan
 * whose trailing permutation passes run as one out-of-place pass (see
 * KATE_PLAN_NO_SCATTER) moves the state between two buffers on every
 * execute(): call this again after each execute() before using the address. */
int kate_sim_device_state(kate_sim* sim, void** dptr);
int kate_sim_last_timings(kate_sim* sim, kate_timings* out);
int kate_sim_synchronize(kate_sim* sim);
/* HIP stream the simulator enqueues on (hipStream_t, as void*). */
int kate_sim_stream(kate_sim* sim, void** stream);
/* kate_sim_run with a HIP event pair around every pass on the simulator's
 * stream; pass_ms[i] (i < cap) receives the duration of pass i and
 * pass_is_remap[i] whether it was a qubit remap (either may be NULL).
 * Returns the number of passes through *npasses.  No counterpart in Kate,
 * whose execute(true) only prints per-group wall times
 * (HIPcode/simulator.hpp:322-351). */
int kate_sim_run_timed(kate_sim* sim, float* pass_ms, int32_t* pass_is_remap, int32_t cap, int32_t* npasses);
/* Calibration for the passes' roofline (no counterpart in Kate): for each
 * pass of the plan, the time of a plain copy of every tile with the pass's
 * own addressing (each amplitude read once and written once, no
 * arithmetic): in place with the pass's tile positions (no LDS), or, for a
 * pass that writes out of place (scatter or relabelled), through LDS and
 * the same store into the second buffer, which does not become the state;
 * i.e. the HBM time the pass's access pattern allows.  The state is left
 * unchanged.  pass_ms[i] for i < cap; *npasses passes. */
int kate_sim_copy_ceiling(kate_sim* sim, float* pass_ms, int32_t cap, int32_t* npasses);
/* kate_sim_execute with a HIP event after every pass on the simulator's
 * stream: pass_ms[i] is pass i's time and pass_tiles[i] the number of tiles
 * it launched (fewer than 2^(n - tile qubits) when the run from |0...0>
 * skips tiles that are known to be zero); either may be NULL.  The passes
 * run are exactly those of kate_sim_execute (*npasses of them).  No
 * counterpart in Kate, whose execute(true) only prints per-group wall times
 * (HIPcode/simulator.hpp:322-351). */
int kate_sim_execute_timed(kate_sim* sim, const double* theta_phi_in, double* theta_phi_out, float* pass_ms,
                           uint64_t* pass_tiles, int32_t cap, int32_t* npasses);
/* Pass plan as text (one line per pass), for inspection and tests. */
int kate_sim_describe(kate_sim* sim, char* buf, size_t len);
/* The same plan without a device (no counterpart in Kate, whose grouping is
 * fixed by gateGrouping): pass and remap counts plus the text form. */
int kate_plan_describe(const kate_program* program, const kate_options* options, int32_t* passes,
                       int32_t* remap_passes, char* buf, size_t len);

/* ---- circuit compiler: backendagnosticcode/Circuit.hpp, QuantumCircuit.hpp ---- */
/* Circuit(int nqbits) (QuantumCircuit.hpp:18-20). */
int kate_circuit_create(int32_t nqubits, kate_circuit** out);
int kate_circuit_destroy(kate_circuit* c);
/* appendGate (QuantumCircuit.hpp:25-37) of Gate(id, qbits, optarg)
 * (basic_host_types.hpp:198-228) or, for id 0, Gate(Matrix, qbits)
 * (:229-245); matrix is 2^nq x 2^nq interleaved complex128, row-major. */
int kate_circuit_append(kate_circuit* c, int32_t id, int32_t nq, const int32_t* qubits, int32_t optarg,
                        const double* matrix);
/* Circuit::gateScheduling (Circuit.hpp:12-98). */
int kate_circuit_schedule(kate_circuit* c);
/* Circuit::gateFusion (Circuit.hpp:99-175) with CPUmergeGate (CPUcode/GateMerger.hpp:8-55). */
int kate_circuit_fuse(kate_circuit* c, int32_t max_qubits, double merge_time_matters);
/* Circuit::gateGrouping (Circuit.hpp:176-206). */
int kate_circuit_group(kate_circuit* c, int32_t group_qubits);
/* Circuit::allocate (Circuit.hpp:207-348). */
int kate_circuit_allocate(kate_circuit* c, int32_t gpu_log2, int32_t max_local);
/* QuantumCircuit::compileDefault (QuantumCircuit.hpp:126-204). */
int kate_circuit_compile_default(kate_circuit* c, int32_t gpu_log2, int32_t max_local);
/* Circuit::compileOPT (Circuit.hpp:349-354). */
int kate_circuit_compile_opt(kate_circuit* c, int32_t fuse_qubits, double merge_time_matters, int32_t group_qubits,
                             int32_t gpu_log2, int32_t max_local);
/* Circuit::exportqcx (Circuit.hpp:763-774). */
int kate_circuit_export_qcx(kate_circuit* c, const char* path);
/* Read back: counts, gate i, groups, instructions, permutations. */
int kate_circuit_info(const kate_circuit* c, int32_t* nqubits, int32_t* ngates, int32_t* ngroups, int32_t* ninstr);
int kate_circuit_gate(const kate_circuit* c, int32_t i, int32_t* id, int32_t* nq, int32_t* qubits, int32_t* optarg,
                      double* matrix /* may be NULL */);
int kate_circuit_groups(const kate_circuit* c, int32_t* ends, uint64_t* masks);
/* Instruction i: kind, and up to cap ints of its pair list; returns the
 * pair-list length through *npairs_ints. */
int kate_circuit_instruction(const kate_circuit* c, int32_t i, int32_t* kind, int32_t* pairs, int32_t cap,
                             int32_t* npairs_ints);
int kate_circuit_permutations(const kate_circuit* c, int32_t* initial, int32_t* final_inverse);
/* Simulator(QuantumCircuit, 1) on this circuit (compiled with
 * compileDefault(0) first if it has no instructions, HIPcode/simulator.hpp:261-264). */
int kate_sim_create_from_circuit(const kate_circuit* c, const kate_options* options, kate_sim** out);
/* kate_plan_describe for this circuit (compiled as kate_sim_create_from_circuit would). */
int kate_circuit_plan(const kate_circuit* c, const kate_options* options, int32_t* passes, int32_t* remap_passes,
                      char* buf, size_t len);

/* ---- batches: many small circuits in one persistent launch ----
 * No counterpart in Kate (one circuit per Simulator).  Each circuit is
 * planned as kate_sim_create would plan it, with one tile shape for the whole
 * batch; a run initialises every state to |0...0>, executes every
 * (circuit, pass, tile) task in one persistent kernel, and reduces <Z_q> for
 * every qubit of every circuit (FP64, fixed order). */
typedef struct kate_batch kate_batch;
int kate_batch_create(const kate_program* programs, int32_t count, const kate_options* options, kate_batch** out);
int kate_batch_create_from_circuits(const kate_circuit* const* circuits, int32_t count, const kate_options* options,
                                    kate_batch** out);
int kate_batch_destroy(kate_batch* batch);
/* Capture the run sequence (kernels only) into a HIP graph; later runs replay it. */
int kate_batch_capture(kate_batch* batch);
/* One run, synchronous; kate_batch_last_ms reports its device time. */
int kate_batch_run(kate_batch* batch);
int kate_batch_info(const kate_batch* batch, int32_t* count, int64_t* z_len, int32_t* tasks, int32_t* passes,
                    int32_t* grid);
/* <Z_q> of every circuit for real qubit q (the circuit's
 * final_inverse_permutation applied), concatenated in circuit order (z_len
 * values); norms (may be NULL) receives each circuit's sum |a|^2. */
int kate_batch_expect_z(kate_batch* batch, double* out, int64_t len, double* norms);
int kate_batch_last_ms(const kate_batch* batch, float* ms);

/* ---- serving: continuous batching of simulate requests ----
 * No counterpart in Kate.  One HBM arena reserved at creation (no allocation
 * while serving), a pool of streams and one scheduler thread: requests are
 * planned off the caller's thread, admitted first come first served as soon
 * as their footprint fits the free arena, and run on the stream pool so that
 * requests in flight overlap on the device.  Each returns <Z_S> for up to 15
 * real-qubit masks (FP64 accumulation, normalised).  A request of 22 qubits
 * or more holds two state buffers in the arena (its passes run out of place,
 * relabelled); a smaller one holds one. */
typedef struct kate_server_options {
    int32_t dtype;
    int32_t device;
    int32_t streams;          /* 0 = 4 */
    int32_t max_active;       /* requests in flight; 0 = 32 */
    int64_t hbm_budget_bytes; /* arena size; 0 = half the free device memory */
    int32_t tile_qubits;      /* as kate_options */
    int32_t low_qubits;
    int32_t use_mfma;
    int32_t micro_bits;       /* as kate_options */
    int32_t reserved[2];
} kate_server_options;

typedef struct kate_request_stats {
    double submit_s;  /* host steady-clock seconds */
    double admit_s;
    double done_s;
    double latency_ms; /* done - submit */
    int32_t passes;
    int32_t nqubits;
    int64_t footprint_bytes;
} kate_request_stats;

typedef struct kate_server kate_server;
int kate_server_create(const kate_server_options* options, kate_server** out);
/* Finishes the requests in flight, then releases everything. */
int kate_server_destroy(kate_server* server);
/* theta_phi: NULL for |0...0>, else the product-state input of
 * kate_sim_init_product.  Non-blocking; *id names the request. */
int kate_server_submit(kate_server* server, const kate_program* program, const double* theta_phi,
                       const uint64_t* zmasks, int32_t nmasks, uint64_t* id);
int kate_server_submit_circuit(kate_server* server, const kate_circuit* circuit, const double* theta_phi,
                               const uint64_t* zmasks, int32_t nmasks, uint64_t* id);
/* Blocks until the request is done; writes min(cap, nmasks) values. */
int kate_server_wait(kate_server* server, uint64_t id, double* zout, int32_t cap, kate_request_stats* stats);
int kate_server_pending(kate_server* server, int32_t* queued, int32_t* active, int64_t* completed,
                        int64_t* largest_free_bytes);

#ifdef __cplusplus
}
#endif
#endif /* KATE_HIP_H */
