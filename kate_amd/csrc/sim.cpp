// sim.cpp -- the kate_sim_* C ABI (include/kate_hip.h).
//
// One simulator = one HIP stream, one device state buffer (hipMalloc of
// 2^n amplitudes, the reference's HIPcode/simulator.hpp:284), and the whole
// program uploaded once as three flat arrays (micro-steps, ops, coefficients)
// instead of the reference's three hipMallocs per gate
// (HIPcode/GPUGate.hpp:237-260).  Running the program enqueues one kernel
// per pass with no host synchronisation in between (the reference
// synchronises and reallocates per group, HIPcode/simulator.hpp:777-809).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kate_hip.h"
#include "kernels.hpp"
#include "plan.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define KATE_HIP(x)                                                                           \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) return fail(KATE_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

bool valid_perm(const int32_t* p, int n)
{
    std::vector<int> seen((size_t)n, 0);
    for (int i = 0; i < n; i++) {
        if (p[i] < 0 || p[i] >= n || seen[(size_t)p[i]]) return false;
        seen[(size_t)p[i]] = 1;
    }
    return true;
}

std::string to_program(const kate_program* in, kate::Program& p)
{
    if (!in) return "null program";
    if (in->nqubits < 1 || in->nqubits > KATE_MAX_QUBITS) return "nqubits out of range";
    if (in->ngates < 0 || (in->ngates > 0 && !in->gates)) return "bad gate array";
    p.nqubits = in->nqubits;
    const int n = in->nqubits;
    for (int g = 0; g < in->ngates; g++) {
        const kate_gate_rec& r = in->gates[g];
        kate::PlanGate pg;
        pg.id = r.id;
        pg.k = r.nq;
        pg.optarg = r.optarg;
        const std::string where = "gate " + std::to_string(g) + ": ";
        switch (r.id) {
        case KATE_H: if (r.nq != 1) return where + "H acts on exactly 1 qubit"; break;
        case KATE_CNOT: if (r.nq != 2) return where + "CNOT acts on exactly 2 qubits"; break;
        case KATE_CRK:
            if (r.nq != 2) return where + "CRk acts on exactly 2 qubits";
            if (r.optarg < 0 || r.optarg > 62) return where + "CRk k out of range";
            break;
        case KATE_TOFFOLI: if (r.nq != 3) return where + "TOFFOLI acts on exactly 3 qubits"; break;
        case KATE_DENSE:
            if (r.nq < 1 || r.nq > 13) return where + "dense gate size out of range";
            if (!in->matrix_pool || r.matrix_offset < 0 ||
                r.matrix_offset + (int64_t(1) << (2 * r.nq)) > in->matrix_pool_len)
                return where + "matrix outside the pool";
            pg.moff = (int64_t)(p.mat.size() / 2);
            p.mat.insert(p.mat.end(), in->matrix_pool + 2 * r.matrix_offset,
                         in->matrix_pool + 2 * (r.matrix_offset + (int64_t(1) << (2 * r.nq))));
            break;
        default:
            return where + "unsupported gate id " + std::to_string(r.id) + " (Kate computes 0, 2, 3, 4, 5)";
        }
        for (int i = 0; i < r.nq; i++) {
            const int q = r.qubits[i];
            if (q < 0 || q >= n) return where + "qubit out of range";
            if (pg.mask & (1ull << q)) return where + "a qubit is present twice";
            pg.q[i] = q;
            pg.mask |= 1ull << q;
        }
        p.gates.push_back(pg);
    }
    if (in->ngroups > 0) {
        if (!in->group_end) return "bad group array";
        int prev = 0;
        for (int i = 0; i < in->ngroups; i++) {
            if (in->group_end[i] < prev || in->group_end[i] > in->ngates) return "group ends out of order";
            prev = in->group_end[i];
            p.group_end.push_back(in->group_end[i]);
            p.group_mask.push_back(in->group_mask ? in->group_mask[i] : 0);
        }
        if (prev != in->ngates) return "groups do not cover every gate";
    } else {
        for (int g = 0; g < in->ngates; g++) {
            p.group_end.push_back(g + 1);
            p.group_mask.push_back(p.gates[(size_t)g].mask);
        }
    }
    if (in->ninstr > 0) {
        if (!in->instr_kind) return "bad instruction array";
        size_t computes = 0;
        for (int i = 0; i < in->ninstr; i++) {
            const int k = in->instr_kind[i];
            if (k != 0 && k != 1) return "instruction kind must be 0 or 1";
            p.instr_kind.push_back(k);
            std::vector<int> pairs;
            if (k == 0) {
                const int np = in->instr_npairs ? in->instr_npairs[i] : 0;
                const int off = in->instr_pair_offset ? in->instr_pair_offset[i] : 0;
                for (int j = 0; j < 2 * np; j++) {
                    const int q = in->pairs[off + j];
                    if (q < 0 || q >= n) return "swap position out of range";
                    pairs.push_back(q);
                }
            } else {
                computes++;
            }
            p.instr_pairs.push_back(pairs);
        }
        if (computes != p.group_end.size()) return "compute instructions do not match the groups";
    } else {
        p.instr_kind.assign(p.group_end.size(), 1);
        p.instr_pairs.assign(p.group_end.size(), std::vector<int>());
    }
    for (int i = 0; i < n; i++) {
        p.initial_permutation.push_back(in->initial_permutation ? in->initial_permutation[i] : i);
        p.final_inverse_permutation.push_back(in->final_inverse_permutation ? in->final_inverse_permutation[i] : i);
    }
    if (!valid_perm(p.initial_permutation.data(), n)) return "initial_permutation is not a permutation";
    if (!valid_perm(p.final_inverse_permutation.data(), n)) return "final_inverse_permutation is not a permutation";
    return "";
}

}  // namespace

namespace kate {
// Shared with the other C-ABI files so kate_last_error() reports any of them.
int set_error(int code, const std::string& msg) { return fail(code, msg); }
std::string program_from_abi(const kate_program* in, Program& p) { return to_program(in, p); }
}  // namespace kate

struct kate_sim {
    int device = 0;
    hipStream_t stream = nullptr;
    int dtype = KATE_C128;
    size_t amp_bytes = 16;
    kate::Program prog;
    kate::Plan plan;
    int nq = 0, nq_int = 0;
    void* d_state = nullptr;
    void* d_state2 = nullptr;  // the other buffer of the scatter pass (which then swaps the two)
    kate::ScatterArgs* d_scatter = nullptr;  // that pass's store offsets
    kate::MicroStep* d_micro = nullptr;
    kate::MicroOp* d_ops = nullptr;
    void* d_coef = nullptr;
    uint32_t* d_raw = nullptr;
    double* d_scratch = nullptr;
    double* d_out = nullptr;
    uint64_t* d_masks = nullptr;
    double* d_factors = nullptr;
    std::vector<double> h_factors;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    kate_timings last{};
    bool support = true;  // execute() from |0...0> skips tiles known to be zero
    bool defer = true;    // execute() measures through the trailing permutation passes
    bool pending = false; // those passes have not run on the state yet
};

namespace {

template <typename T>
int upload(T** dst, const std::vector<T>& src, hipStream_t s)
{
    const size_t bytes = std::max<size_t>(1, src.size()) * sizeof(T);
    KATE_HIP(hipMalloc((void**)dst, bytes));
    // synchronous: the source may be a temporary
    if (!src.empty()) KATE_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    (void)s;
    return KATE_OK;
}

void release(kate_sim* s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    (void)hipFree(s->d_state);
    (void)hipFree(s->d_state2);
    (void)hipFree(s->d_scatter);
    (void)hipFree(s->d_micro);
    (void)hipFree(s->d_ops);
    (void)hipFree(s->d_coef);
    (void)hipFree(s->d_raw);
    (void)hipFree(s->d_scratch);
    (void)hipFree(s->d_out);
    (void)hipFree(s->d_masks);
    (void)hipFree(s->d_factors);
    for (auto& e : s->ev)
        if (e) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

// Runs every pass on the current state, or, with `from_zero`, on |0...0>
// without an init kernel and without reading it (kate::zero_start_pass:
// the first pass makes its tile in registers and, unless the support
// skipping is off, the passes follow the support of the state).
// Passes [begin, end) -- all of them by default.
// With `after` (events, one per pass in [begin, end)) an event is recorded
// after each pass and `tiles` receives each pass's launched tile count.
// One pass on the state; the scatter pass writes the other buffer, which
// then holds the state.
int launch_plan_pass(kate_sim* s, const kate::PassArgs& pa, const kate::DevPlan& dp)
{
    if (!pa.scatter) return kate::launch_pass(s->dtype, s->plan.tile_bits, s->d_state, s->nq_int, pa, dp, s->stream);
    if (!s->d_state2 || !s->d_scatter) return (int)hipErrorInvalidValue;
    int e;
    if (pa.micro_count) {  // a gate pass that stores through the trailing permutation
        kate::DevPlan d = dp;
        d.dst = s->d_state2;
        d.scatter = s->d_scatter;
        e = kate::launch_pass(s->dtype, s->plan.tile_bits, s->d_state, s->nq_int, pa, d, s->stream);
    } else {
        e = kate::launch_scatter(s->dtype, s->plan.tile_bits, s->d_state, s->d_state2, pa, s->d_scatter, s->stream);
    }
    if (!e) std::swap(s->d_state, s->d_state2);
    return e;
}

int run_passes(kate_sim* s, bool from_zero = false, size_t begin = 0, size_t end = (size_t)-1,
               hipEvent_t* after = nullptr, uint64_t* tiles = nullptr)
{
    kate::DevPlan dp{s->d_micro, s->d_ops, s->d_coef, s->d_raw};
    end = std::min(end, s->plan.passes.size());
    for (size_t i = begin; i < end; i++) {
        const kate::PassArgs pa = from_zero ? kate::zero_start_pass(s->plan, i, s->support, end) : s->plan.passes[i];
        int e = launch_plan_pass(s, pa, dp);
        if (e) return fail(KATE_ERR_HIP, std::string("pass launch: ") + hipGetErrorString((hipError_t)e));
        if (tiles) tiles[i - begin] = 1ull << pa.nrest;
        if (after) KATE_HIP(hipEventRecord(after[i - begin], s->stream));
    }
    return KATE_OK;
}

// execute() can leave the trailing permutation passes unrun (its
// measurement reads through them); every other use of the state runs them
// first.
int ensure_final(kate_sim* s)
{
    if (!s->pending) return KATE_OK;
    s->pending = false;
    return run_passes(s, false, s->plan.trail_begin);
}

int init_product(kate_sim* s, const double* tp)
{
    // factor[b*nq_int + j] for virtual position j (Kate places real qubit i at
    // initial_permutation[i], HIPcode/simulator.hpp:452-458); padding is |0>.
    const int n = s->nq, m = s->nq_int;
    std::vector<double> f(4 * (size_t)m, 0.0);
    for (int j = 0; j < m; j++) f[2 * (size_t)j] = 1.0;
    const double pi = 3.1415926535897932384626433;
    for (int i = 0; i < n; i++) {
        const double th = tp[2 * i], ph = tp[2 * i + 1];
        const int v = s->prog.initial_permutation[(size_t)i];
        const double sn = std::sin(th * pi / 2);
        f[2 * (size_t)v] = std::cos(th * pi / 2);
        f[2 * (size_t)v + 1] = 0.0;
        f[2 * ((size_t)v + m)] = std::cos(ph) * sn;
        f[2 * ((size_t)v + m) + 1] = std::sin(ph) * sn;
    }
    KATE_HIP(hipStreamSynchronize(s->stream));  // the previous factors may still be in use
    s->h_factors.swap(f);                        // stays alive until the next product-state init
    KATE_HIP(hipMemcpyAsync(s->d_factors, s->h_factors.data(), s->h_factors.size() * sizeof(double),
                            hipMemcpyHostToDevice, s->stream));
    int e = kate::launch_init_product(s->dtype, s->d_state, s->nq_int, s->d_factors, s->stream);
    if (e) return fail(KATE_ERR_HIP, std::string("init launch: ") + hipGetErrorString((hipError_t)e));
    return KATE_OK;
}

// Kate's proba_state from the amplitude sums (HIPcode/simulator.hpp:519-536).
int measure(kate_sim* s, double* out)
{
    // with the trailing passes pending, amplitude i is still at offset
    // trail_inv(i)
    const uint64_t* gather = s->pending ? s->plan.trail_inv.data() : nullptr;
    int e = kate::launch_amp_sums(s->dtype, s->d_state, s->nq_int, s->d_scratch, s->d_out, s->stream, gather);
    if (e) return fail(KATE_ERR_HIP, std::string("measure launch: ") + hipGetErrorString((hipError_t)e));
    std::vector<double> sums(2 + 2 * (size_t)s->nq_int);
    KATE_HIP(hipMemcpyAsync(sums.data(), s->d_out, sums.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    KATE_HIP(hipStreamSynchronize(s->stream));
    const int n = s->nq;
    const double den = std::pow(1.4142135623730951, (double)n);
    const double pi = 3.1415926535897932384626433;
    for (int q = 0; q < n; q++) {
        const double s1r = sums[2 + 2 * (size_t)q], s1i = sums[3 + 2 * (size_t)q];
        const double s0r = sums[0] - s1r, s0i = sums[1] - s1i;
        const double nn = den * den;
        const double a0 = (s0r * den) / nn, b0 = (den * s0i) / nn;
        const double a1 = (s1r * den) / nn, b1 = (den * s1i) / nn;
        const double n0 = std::sqrt(a0 * a0 + b0 * b0), n1 = std::sqrt(a1 * a1 + b1 * b1);
        double th, ph;
        if (n0 < 0.0000000000001) {
            th = 1;
            ph = 0;
        } else if (n1 < 0.00000000000001) {
            th = 0;
            ph = 0;
        } else {
            th = std::atan(n1 / n0) / (pi / 2);
            const double d = a0 * a0 + b0 * b0;
            ph = std::atan(((a0 * b1 - a1 * b0) / d) / ((a1 * a0 + b1 * b0) / d));
        }
        const int r = s->prog.final_inverse_permutation[(size_t)q];
        out[2 * r] = th;
        out[2 * r + 1] = ph;
    }
    return KATE_OK;
}

int zsums(kate_sim* s, const std::vector<uint64_t>& masks, std::vector<double>& out)
{
    out.assign(masks.size(), 0.0);
    for (size_t b = 0; b < masks.size(); b += 16) {
        const int cnt = (int)std::min<size_t>(16, masks.size() - b);
        KATE_HIP(hipMemcpyAsync(s->d_masks, masks.data() + b, cnt * sizeof(uint64_t), hipMemcpyHostToDevice, s->stream));
        int e = kate::launch_zsums(s->dtype, s->d_state, s->nq_int, s->d_masks, cnt, s->d_scratch, s->d_out, s->stream);
        if (e) return fail(KATE_ERR_HIP, std::string("zsum launch: ") + hipGetErrorString((hipError_t)e));
        double h[16];
        KATE_HIP(hipMemcpyAsync(h, s->d_out, sizeof(h), hipMemcpyDeviceToHost, s->stream));
        KATE_HIP(hipStreamSynchronize(s->stream));
        for (int j = 0; j < cnt; j++) out[b + (size_t)j] = h[j];
    }
    return KATE_OK;
}

// Real-qubit mask -> virtual-position mask.
uint64_t to_virtual(const kate_sim* s, uint64_t real)
{
    uint64_t v = 0;
    for (int p = 0; p < s->nq; p++)
        if ((real >> s->prog.final_inverse_permutation[(size_t)p]) & 1ull) v |= 1ull << p;
    return v;
}

}  // namespace

namespace kate {
std::string plan_options(const kate_options& opt, kate::PlanOptions& po)
{
    if (opt.dtype != KATE_C128 && opt.dtype != KATE_C64) return "dtype must be KATE_C128 or KATE_C64";
    // complex128: 12-qubit tiles (64 KiB of LDS) by default; complex64: the
    // planner picks 13 (64 KiB) or 12 (32 KiB) per circuit
    po.tile_bits = opt.tile_qubits ? opt.tile_qubits : (opt.dtype == KATE_C128 ? 12 : 13);
    po.auto_tile = opt.tile_qubits == 0 && opt.dtype == KATE_C64;
    po.low_qubits = opt.low_qubits ? opt.low_qubits : (opt.dtype == KATE_C128 ? 3 : 4);
    po.auto_low = opt.low_qubits == 0;  // 0: the planner picks it per circuit
    // the reductions read whole 4096-amplitude blocks
    po.min_qubits = std::max(po.tile_bits, 12);
    if (opt.use_mfma && opt.dtype != KATE_C64) return "use_mfma needs KATE_C64 (the FP32 matrix-core path)";
    if (opt.use_mfma < 0 || opt.use_mfma > 3) return "use_mfma must be 0..3";
    po.mfma = opt.use_mfma >= 1;
    po.mfma_fuse = opt.use_mfma >= 2;
    po.mfma_bf16 = opt.use_mfma == 3;
    po.micro_bits = opt.micro_bits ? opt.micro_bits : 4;
    po.amp_bytes = opt.dtype == KATE_C128 ? 16 : 8;
    if (opt.plan_flags &
        ~(KATE_PLAN_NO_EULER | KATE_PLAN_NO_DIRECT | KATE_PLAN_NO_SUPPORT | KATE_PLAN_NO_DEFER | KATE_PLAN_NO_SPLIT |
          KATE_PLAN_NO_REORDER | KATE_PLAN_NO_FACTOR | KATE_PLAN_NO_SCATTER | KATE_PLAN_NO_RELABEL |
          KATE_PLAN_NO_MATRIX5))
        return "unknown plan_flags bits";
    po.split = !(opt.plan_flags & KATE_PLAN_NO_SPLIT);
    po.reorder = !(opt.plan_flags & KATE_PLAN_NO_REORDER);
    po.defer = !(opt.plan_flags & KATE_PLAN_NO_DEFER);
    po.euler = !(opt.plan_flags & KATE_PLAN_NO_EULER);
    po.factor = !(opt.plan_flags & KATE_PLAN_NO_FACTOR);
    po.scatter = !(opt.plan_flags & KATE_PLAN_NO_SCATTER);
    po.relabel = !(opt.plan_flags & KATE_PLAN_NO_RELABEL);
    po.direct = !(opt.plan_flags & KATE_PLAN_NO_DIRECT);
    po.support = !(opt.plan_flags & KATE_PLAN_NO_SUPPORT);
    // complex64: dense 5-qubit gates on the matrix cores even without
    // use_mfma (28 qubits, 60 Haar 5-qubit gates: 62 ms against 84 ms on the
    // VALU; 4-qubit gates stay on the register path, 24 against 31 ms)
    po.mfma5 = opt.dtype == KATE_C64 && po.micro_bits == 4 && !(opt.plan_flags & KATE_PLAN_NO_MATRIX5);
    if (po.micro_bits != 3 && po.micro_bits != 4) return "micro_bits must be 3 or 4";
    if (po.micro_bits == 3 && opt.dtype != KATE_C128) return "micro_bits 3 needs KATE_C128";
    if (po.micro_bits == 3 && opt.use_mfma) return "micro_bits 3 excludes use_mfma";
    if (po.tile_bits < 10 || po.tile_bits > 13) return "tile_qubits must be in 10..13";
    if (po.low_qubits < 0 || po.low_qubits > po.tile_bits) return "low_qubits out of range";
    if (opt.batch_wgs_per_cu < 0 || opt.batch_wgs_per_cu > 32) return "batch_wgs_per_cu must be in 0..32";
    return "";
}
}  // namespace kate
using kate::plan_options;

extern "C" {

const char* kate_status_string(int st)
{
    switch (st) {
    case KATE_OK: return "ok";
    case KATE_ERR_ARG: return "invalid argument";
    case KATE_ERR_NO_DEVICE: return "no GPU";
    case KATE_ERR_HIP: return "HIP error";
    case KATE_ERR_OOM: return "out of device memory";
    case KATE_ERR_PROGRAM: return "invalid program";
    case KATE_ERR_UNSUPPORTED: return "unsupported";
    case KATE_ERR_BUSY: return "busy";
    default: return "unknown";
    }
}

const char* kate_last_error(void) { return g_err.c_str(); }

int kate_device_count(int* count)
{
    if (!count) return fail(KATE_ERR_ARG, "null count");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return fail(KATE_ERR_NO_DEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    return KATE_OK;
}

int kate_device_info(char* buf, size_t len)
{
    int n = 0;
    int rc = kate_device_count(&n);
    if (rc) return rc;
    std::string s;
    for (int d = 0; d < n; d++) {
        hipDeviceProp_t p;
        KATE_HIP(hipGetDeviceProperties(&p, d));
        char line[512];
        snprintf(line, sizeof(line),
                 "device %d: %s (%s) CUs=%d HBM=%.1f GiB LDS/workgroup=%zu KiB LDS/CU=%zu KiB clock=%d MHz\n", d, p.name,
                 p.gcnArchName, p.multiProcessorCount, p.totalGlobalMem / 1073741824.0, p.sharedMemPerBlock / 1024,
                 (size_t)p.maxSharedMemoryPerMultiProcessor / 1024, p.clockRate / 1000);
        s += line;
    }
    if (buf && len) {
        std::strncpy(buf, s.c_str(), len - 1);
        buf[len - 1] = 0;
    }
    return KATE_OK;
}


int kate_plan_describe(const kate_program* program, const kate_options* options, int32_t* passes,
                       int32_t* remap_passes, char* buf, size_t len)
{
    kate_options opt{};
    if (options) opt = *options;
    kate::PlanOptions po;
    std::string err = plan_options(opt, po);
    if (!err.empty()) return fail(KATE_ERR_ARG, err);
    kate::Program prog;
    err = to_program(program, prog);
    if (!err.empty()) return fail(KATE_ERR_PROGRAM, err);
    kate::Plan plan;
    err = kate::build_plan(prog, po, plan);
    if (!err.empty()) return fail(KATE_ERR_PROGRAM, "planning: " + err);
    if (passes) *passes = plan.gate_passes;
    if (remap_passes) *remap_passes = plan.remap_passes;
    if (buf && len) {
        // KATE_DESCRIBE_OPS=1 adds every micro-step and its register ops.
        const char* v = std::getenv("KATE_DESCRIBE_OPS");
        std::string d = plan.describe(v && v[0] == '1');
        std::strncpy(buf, d.c_str(), len - 1);
        buf[len - 1] = 0;
    }
    return KATE_OK;
}

int kate_sim_create(const kate_program* program, const kate_options* options, kate_sim** out)
{
    if (!out) return fail(KATE_ERR_ARG, "null output handle");
    *out = nullptr;
    kate_options opt{};
    if (options) opt = *options;
    if (opt.dtype != KATE_C128 && opt.dtype != KATE_C64) return fail(KATE_ERR_ARG, "dtype must be KATE_C128 or KATE_C64");
    int count = 0;
    if (kate_device_count(&count) != KATE_OK || count == 0)
        return fail(KATE_ERR_NO_DEVICE, "no GPU visible: the library has no CPU path");
    if (opt.device < 0 || opt.device >= count) return fail(KATE_ERR_ARG, "device ordinal out of range");

    kate_sim* s = new kate_sim();
    s->device = opt.device;
    s->dtype = opt.dtype;
    s->amp_bytes = opt.dtype == KATE_C128 ? 16 : 8;
    std::string err = to_program(program, s->prog);
    if (!err.empty()) {
        delete s;
        return fail(KATE_ERR_PROGRAM, err);
    }
    kate::PlanOptions po;
    err = plan_options(opt, po);
    if (!err.empty()) {
        delete s;
        return fail(KATE_ERR_ARG, err);
    }
    s->support = po.support;
    s->defer = po.defer;
    if (po.scatter) {
        // the scatter pass needs a second state buffer: only when the GPU
        // has room for both and 1 GiB to spare
        size_t fr = 0, tot = 0;
        const int nmax = std::max(s->prog.nqubits, std::max(po.min_qubits, 13));
        const unsigned __int128 need = ((unsigned __int128)2 << nmax) * s->amp_bytes + (1ull << 30);
        if (nmax >= 60 || hipSetDevice(s->device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess ||
            (unsigned __int128)fr < need)
            po.scatter = false;
    }
    err = kate::build_plan(s->prog, po, s->plan);
    if (!err.empty()) {
        delete s;
        return fail(KATE_ERR_PROGRAM, "planning: " + err);
    }
    s->nq = s->prog.nqubits;
    s->nq_int = s->plan.nq_int;

    hipError_t he = hipSetDevice(s->device);
    if (he != hipSuccess) {
        delete s;
        return fail(KATE_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(he));
    }
    int rc = KATE_OK;
    auto step = [&](hipError_t e, const char* what) {
        if (rc == KATE_OK && e != hipSuccess)
            rc = fail(e == hipErrorOutOfMemory ? KATE_ERR_OOM : KATE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    step(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : s->ev) step(hipEventCreate(&e), "hipEventCreate");
    step(hipMalloc(&s->d_state, (1ull << s->nq_int) * s->amp_bytes), "hipMalloc(state)");
    if (!s->plan.scatter_args.empty()) {
        const size_t sb = s->plan.scatter_args.size() * sizeof(kate::ScatterArgs);
        step(hipMalloc(&s->d_state2, (1ull << s->nq_int) * s->amp_bytes), "hipMalloc(state 2)");
        step(hipMalloc((void**)&s->d_scatter, sb), "hipMalloc(scatter)");
        if (rc == KATE_OK)
            step(hipMemcpy(s->d_scatter, s->plan.scatter_args.data(), sb, hipMemcpyHostToDevice), "upload(scatter)");
    }
    step(hipMalloc((void**)&s->d_scratch, kate::reduce_scratch_doubles(s->nq_int) * sizeof(double)), "hipMalloc(scratch)");
    step(hipMalloc((void**)&s->d_out, (2 + 2 * (size_t)s->nq_int + 64) * sizeof(double)), "hipMalloc(out)");
    step(hipMalloc((void**)&s->d_masks, 16 * sizeof(uint64_t)), "hipMalloc(masks)");
    step(hipMalloc((void**)&s->d_factors, 4 * (size_t)s->nq_int * sizeof(double)), "hipMalloc(factors)");
    if (rc == KATE_OK) rc = upload(&s->d_micro, s->plan.micro, s->stream);
    if (rc == KATE_OK) rc = upload(&s->d_ops, s->plan.ops, s->stream);
    if (rc == KATE_OK) rc = upload(&s->d_raw, s->plan.raw, s->stream);
    if (rc == KATE_OK) {
        if (s->dtype == KATE_C128) {
            rc = upload((double**)&s->d_coef, s->plan.coef, s->stream);
        } else {
            std::vector<float> c(s->plan.coef.begin(), s->plan.coef.end());
            rc = upload((float**)&s->d_coef, c, s->stream);
        }
    }
    if (rc == KATE_OK) step(hipStreamSynchronize(s->stream), "upload");
    if (rc != KATE_OK) {
        std::string keep = g_err;
        release(s);
        g_err = keep;
        return rc;
    }
    s->last.passes = s->plan.gate_passes;
    s->last.remap_passes = s->plan.remap_passes;
    s->last.micro_steps = (int32_t)s->plan.micro.size();
    s->last.gates = (int32_t)s->prog.gates.size();
    s->last.bytes_per_pass = 2.0 * (double)(1ull << s->nq_int) * (double)s->amp_bytes;
    *out = s;
    return KATE_OK;
}

int kate_sim_destroy(kate_sim* s)
{
    release(s);
    return KATE_OK;
}

int kate_sim_init_basis(kate_sim* s, uint64_t index)
{
    if (!s) return fail(KATE_ERR_ARG, "null simulator");
    if (index >= (1ull << s->nq)) return fail(KATE_ERR_ARG, "basis index out of range");
    KATE_HIP(hipSetDevice(s->device));
    s->pending = false;  // the whole state is replaced
    int e = kate::launch_init_basis(s->dtype, s->d_state, s->nq_int, index, s->stream);
    if (e) return fail(KATE_ERR_HIP, std::string("init launch: ") + hipGetErrorString((hipError_t)e));
    return KATE_OK;
}

int kate_sim_init_product(kate_sim* s, const double* tp)
{
    if (!s || !tp) return fail(KATE_ERR_ARG, "null argument");
    KATE_HIP(hipSetDevice(s->device));
    s->pending = false;  // the whole state is replaced
    return init_product(s, tp);
}

int kate_sim_run(kate_sim* s)
{
    if (!s) return fail(KATE_ERR_ARG, "null simulator");
    KATE_HIP(hipSetDevice(s->device));
    int rc = ensure_final(s);
    if (rc) return rc;
    return run_passes(s);
}

int kate_sim_run_timed(kate_sim* s, float* pass_ms, int32_t* pass_is_remap, int32_t cap, int32_t* npasses)
{
    if (!s) return fail(KATE_ERR_ARG, "null simulator");
    KATE_HIP(hipSetDevice(s->device));
    if (int rc0 = ensure_final(s)) return rc0;
    const size_t np = s->plan.passes.size();
    if (npasses) *npasses = (int32_t)np;
    std::vector<hipEvent_t> ev(np + 1, nullptr);
    int rc = KATE_OK;
    for (auto& e : ev)
        if (rc == KATE_OK && hipEventCreate(&e) != hipSuccess) rc = fail(KATE_ERR_HIP, "hipEventCreate");
    kate::DevPlan dp{s->d_micro, s->d_ops, s->d_coef, s->d_raw};
    if (rc == KATE_OK && hipEventRecord(ev[0], s->stream) != hipSuccess) rc = fail(KATE_ERR_HIP, "hipEventRecord");
    for (size_t i = 0; i < np && rc == KATE_OK; i++) {
        int e = launch_plan_pass(s, s->plan.passes[i], dp);
        if (e) rc = fail(KATE_ERR_HIP, std::string("pass launch: ") + hipGetErrorString((hipError_t)e));
        else if (hipEventRecord(ev[i + 1], s->stream) != hipSuccess) rc = fail(KATE_ERR_HIP, "hipEventRecord");
    }
    if (rc == KATE_OK && hipEventSynchronize(ev[np]) != hipSuccess) rc = fail(KATE_ERR_HIP, "hipEventSynchronize");
    for (size_t i = 0; i < np && rc == KATE_OK; i++) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess) rc = fail(KATE_ERR_HIP, "hipEventElapsedTime");
        if ((int32_t)i < cap) {
            if (pass_ms) pass_ms[i] = ms;
            if (pass_is_remap) pass_is_remap[i] = s->plan.pass_is_remap[i];
        }
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int kate_sim_copy_ceiling(kate_sim* s, float* pass_ms, int32_t cap, int32_t* npasses)
{
    if (!s || cap < 0) return fail(KATE_ERR_ARG, "bad arguments");
    KATE_HIP(hipSetDevice(s->device));
    if (int rc0 = ensure_final(s)) return rc0;
    const size_t np = s->plan.passes.size();
    if (npasses) *npasses = (int32_t)np;
    std::vector<hipEvent_t> ev(np + 1, nullptr);
    int rc = KATE_OK;
    for (auto& e : ev)
        if (rc == KATE_OK && hipEventCreate(&e) != hipSuccess) rc = fail(KATE_ERR_HIP, "hipEventCreate");
    if (rc == KATE_OK && hipEventRecord(ev[0], s->stream) != hipSuccess) rc = fail(KATE_ERR_HIP, "hipEventRecord");
    for (size_t i = 0; i < np && rc == KATE_OK; i++) {
        // a scatter pass is timed as a plain copy through its own store, into
        // the other buffer, which does not become the state
        int e = s->plan.passes[i].scatter
                    ? kate::launch_scatter(s->dtype, s->plan.tile_bits, s->d_state, s->d_state2, s->plan.passes[i],
                                           s->d_scatter, s->stream)
                    : kate::launch_copy_tiles(s->dtype, s->plan.tile_bits, s->d_state, s->nq_int, s->plan.passes[i],
                                              s->stream);
        if (e) rc = fail(KATE_ERR_HIP, std::string("copy launch: ") + hipGetErrorString((hipError_t)e));
        else if (hipEventRecord(ev[i + 1], s->stream) != hipSuccess) rc = fail(KATE_ERR_HIP, "hipEventRecord");
    }
    if (rc == KATE_OK && hipEventSynchronize(ev[np]) != hipSuccess) rc = fail(KATE_ERR_HIP, "hipEventSynchronize");
    for (size_t i = 0; i < np && rc == KATE_OK; i++) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess) rc = fail(KATE_ERR_HIP, "hipEventElapsedTime");
        if ((int32_t)i < cap && pass_ms) pass_ms[i] = ms;
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int kate_sim_measure_proba(kate_sim* s, double* tp)
{
    if (!s || !tp) return fail(KATE_ERR_ARG, "null argument");
    KATE_HIP(hipSetDevice(s->device));
    return measure(s, tp);  // reads through pending trailing passes
}

namespace {

// kate_sim_execute, optionally with an event after every pass (pass_ev,
// one per pass the run launches: execute_pass_count of them).
size_t execute_pass_count(const kate_sim* s)
{
    return (s->defer && s->plan.trail_begin > 0) ? s->plan.trail_begin : s->plan.passes.size();
}

int execute_impl(kate_sim* s, const double* tp_in, double* tp_out, hipEvent_t* pass_ev, uint64_t* pass_tiles)
{
    KATE_HIP(hipEventRecord(s->ev[0], s->stream));
    // |0...0> is made by the first pass itself (no init pass over HBM) unless
    // there is no pass to make it
    s->pending = false;  // the state is made afresh
    const bool fused_init = !tp_in && !s->plan.passes.empty();
    int rc = KATE_OK;
    if (tp_in) rc = init_product(s, tp_in);
    else if (!fused_init) rc = kate::launch_init_basis(s->dtype, s->d_state, s->nq_int, 0, s->stream);
    if (rc) return rc > 0 ? fail(KATE_ERR_HIP, "init launch failed") : rc;
    KATE_HIP(hipEventRecord(s->ev[1], s->stream));
    // trailing permutation passes are left for whoever reads the state next;
    // the measurement below reads through them
    const size_t n_run = execute_pass_count(s);
    rc = run_passes(s, fused_init, 0, n_run, pass_ev, pass_tiles);
    if (rc) return rc;
    s->pending = n_run < s->plan.passes.size();
    KATE_HIP(hipEventRecord(s->ev[2], s->stream));
    std::vector<double> tmp(2 * (size_t)s->nq);
    rc = measure(s, tp_out ? tp_out : tmp.data());
    if (rc) return rc;
    KATE_HIP(hipEventRecord(s->ev[3], s->stream));
    KATE_HIP(hipEventSynchronize(s->ev[3]));
    float a = 0, b = 0, c = 0;
    KATE_HIP(hipEventElapsedTime(&a, s->ev[0], s->ev[1]));
    KATE_HIP(hipEventElapsedTime(&b, s->ev[1], s->ev[2]));
    KATE_HIP(hipEventElapsedTime(&c, s->ev[2], s->ev[3]));
    s->last.init_ms = a;
    s->last.compute_ms = b;
    s->last.measure_ms = c;
    return KATE_OK;
}

}  // namespace

int kate_sim_execute(kate_sim* s, const double* tp_in, double* tp_out)
{
    if (!s) return fail(KATE_ERR_ARG, "null simulator");
    KATE_HIP(hipSetDevice(s->device));
    return execute_impl(s, tp_in, tp_out, nullptr, nullptr);
}

int kate_sim_execute_timed(kate_sim* s, const double* tp_in, double* tp_out, float* pass_ms, uint64_t* pass_tiles,
                           int32_t cap, int32_t* npasses)
{
    if (!s || cap < 0) return fail(KATE_ERR_ARG, "bad arguments");
    KATE_HIP(hipSetDevice(s->device));
    const size_t np = execute_pass_count(s);
    if (npasses) *npasses = (int32_t)np;
    std::vector<hipEvent_t> ev(np, nullptr);
    std::vector<uint64_t> tiles(np, 0);
    int rc = KATE_OK;
    for (auto& e : ev)
        if (rc == KATE_OK && hipEventCreate(&e) != hipSuccess) rc = fail(KATE_ERR_HIP, "hipEventCreate");
    if (rc == KATE_OK) rc = execute_impl(s, tp_in, tp_out, ev.data(), tiles.data());
    // pass i runs between the end of the init (ev[1]) or of pass i-1 and ev[i]
    for (size_t i = 0; i < np && rc == KATE_OK; i++) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, i ? ev[i - 1] : s->ev[1], ev[i]) != hipSuccess)
            rc = fail(KATE_ERR_HIP, "hipEventElapsedTime");
        if ((int32_t)i < cap) {
            if (pass_ms) pass_ms[i] = ms;
            if (pass_tiles) pass_tiles[i] = tiles[i];
        }
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int kate_sim_expect_z(kate_sim* s, const uint64_t* masks, int32_t count, double* out)
{
    if (!s || (count > 0 && (!masks || !out)) || count < 0) return fail(KATE_ERR_ARG, "bad arguments");
    KATE_HIP(hipSetDevice(s->device));
    if (int rc0 = ensure_final(s)) return rc0;
    std::vector<uint64_t> vm;
    vm.push_back(0);  // norm
    for (int j = 0; j < count; j++) {
        if (s->nq < 64 && (masks[j] >> s->nq)) return fail(KATE_ERR_ARG, "mask names a qubit outside the circuit");
        vm.push_back(to_virtual(s, masks[j]));
    }
    std::vector<double> r;
    int rc = zsums(s, vm, r);
    if (rc) return rc;
    for (int j = 0; j < count; j++) out[j] = r[(size_t)j + 1] / r[0];
    return KATE_OK;
}

int kate_sim_norm(kate_sim* s, double* out)
{
    if (!s || !out) return fail(KATE_ERR_ARG, "null argument");
    KATE_HIP(hipSetDevice(s->device));
    if (int rc0 = ensure_final(s)) return rc0;
    std::vector<double> r;
    int rc = zsums(s, std::vector<uint64_t>{0}, r);
    if (rc) return rc;
    *out = r[0];
    return KATE_OK;
}

int kate_sim_renormalize(kate_sim* s)
{
    double n = 0;
    int rc = kate_sim_norm(s, &n);
    if (rc) return rc;
    if (!(n > 0)) return fail(KATE_ERR_ARG, "state has zero norm");
    int e = kate::launch_scale(s->dtype, s->d_state, s->nq_int, 1.0 / std::sqrt(n), s->stream);
    if (e) return fail(KATE_ERR_HIP, std::string("scale launch: ") + hipGetErrorString((hipError_t)e));
    return KATE_OK;
}

int kate_sim_read_state(kate_sim* s, void* host, uint64_t bytes, int32_t real_order)
{
    if (!s || !host) return fail(KATE_ERR_ARG, "null argument");
    const uint64_t amps = 1ull << s->nq;
    if (bytes != amps * s->amp_bytes) return fail(KATE_ERR_ARG, "buffer size must be 2^n amplitudes");
    KATE_HIP(hipSetDevice(s->device));
    if (int rc0 = ensure_final(s)) return rc0;
    bool identity = true;
    for (int p = 0; p < s->nq; p++)
        if (s->prog.final_inverse_permutation[(size_t)p] != p) identity = false;
    if (!real_order || identity) {
        KATE_HIP(hipMemcpyAsync(host, s->d_state, bytes, hipMemcpyDeviceToHost, s->stream));
        KATE_HIP(hipStreamSynchronize(s->stream));
        return KATE_OK;
    }
    std::vector<unsigned char> tmp(bytes);
    KATE_HIP(hipMemcpyAsync(tmp.data(), s->d_state, bytes, hipMemcpyDeviceToHost, s->stream));
    KATE_HIP(hipStreamSynchronize(s->stream));
    unsigned char* dst = (unsigned char*)host;
    for (uint64_t v = 0; v < amps; v++) {
        uint64_t r = 0;
        for (int p = 0; p < s->nq; p++)
            if ((v >> p) & 1ull) r |= 1ull << s->prog.final_inverse_permutation[(size_t)p];
        std::memcpy(dst + r * s->amp_bytes, tmp.data() + v * s->amp_bytes, s->amp_bytes);
    }
    return KATE_OK;
}

int kate_sim_write_state(kate_sim* s, const void* host, uint64_t bytes)
{
    if (!s || !host) return fail(KATE_ERR_ARG, "null argument");
    const uint64_t amps = 1ull << s->nq;
    if (bytes != amps * s->amp_bytes) return fail(KATE_ERR_ARG, "buffer size must be 2^n amplitudes");
    KATE_HIP(hipSetDevice(s->device));
    s->pending = false;  // the whole state is replaced
    if (s->nq_int > s->nq) {
        int e = kate::launch_init_basis(s->dtype, s->d_state, s->nq_int, 1ull << s->nq_int, s->stream);  // zero fill
        if (e) return fail(KATE_ERR_HIP, "zero fill launch failed");
    }
    KATE_HIP(hipMemcpyAsync(s->d_state, host, bytes, hipMemcpyHostToDevice, s->stream));
    KATE_HIP(hipStreamSynchronize(s->stream));
    return KATE_OK;
}

int kate_sim_overlap(kate_sim* a, kate_sim* b, double* out5)
{
    if (!a || !b || !out5) return fail(KATE_ERR_ARG, "null argument");
    if (a->nq != b->nq || a->nq_int != b->nq_int || a->device != b->device)
        return fail(KATE_ERR_ARG, "overlap needs two simulators of the same size on the same device");
    if (a->prog.final_inverse_permutation != b->prog.final_inverse_permutation)
        return fail(KATE_ERR_ARG, "overlap needs the same virtual qubit order in both states");
    KATE_HIP(hipSetDevice(a->device));
    if (int rc0 = ensure_final(a)) return rc0;
    if (int rc0 = ensure_final(b)) return rc0;
    KATE_HIP(hipStreamSynchronize(b->stream));
    int e = kate::launch_overlap(a->dtype, a->d_state, b->dtype, b->d_state, a->nq_int, a->d_scratch, a->d_out,
                                 a->stream);
    if (e) return fail(KATE_ERR_HIP, std::string("overlap launch: ") + hipGetErrorString((hipError_t)e));
    KATE_HIP(hipMemcpyAsync(out5, a->d_out, 5 * sizeof(double), hipMemcpyDeviceToHost, a->stream));
    KATE_HIP(hipStreamSynchronize(a->stream));
    return KATE_OK;
}

int kate_sim_device_state(kate_sim* s, void** dptr)
{
    if (!s || !dptr) return fail(KATE_ERR_ARG, "null argument");
    KATE_HIP(hipSetDevice(s->device));
    if (int rc0 = ensure_final(s)) return rc0;
    *dptr = s->d_state;
    return KATE_OK;
}

int kate_sim_last_timings(kate_sim* s, kate_timings* out)
{
    if (!s || !out) return fail(KATE_ERR_ARG, "null argument");
    *out = s->last;
    return KATE_OK;
}

int kate_sim_synchronize(kate_sim* s)
{
    if (!s) return fail(KATE_ERR_ARG, "null simulator");
    KATE_HIP(hipStreamSynchronize(s->stream));
    return KATE_OK;
}

int kate_sim_stream(kate_sim* s, void** stream)
{
    if (!s || !stream) return fail(KATE_ERR_ARG, "null argument");
    *stream = (void*)s->stream;
    return KATE_OK;
}

int kate_sim_describe(kate_sim* s, char* buf, size_t len)
{
    if (!s || !buf || !len) return fail(KATE_ERR_ARG, "null argument");
    std::string d = s->plan.describe();
    std::strncpy(buf, d.c_str(), len - 1);
    buf[len - 1] = 0;
    return d.size() < len ? KATE_OK : fail(KATE_ERR_ARG, "buffer too small");
}

}  // extern "C"
