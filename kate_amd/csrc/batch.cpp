// batch.cpp -- many small circuits in one persistent launch (kate_batch_* C ABI).
//
// Every circuit is planned exactly as a single simulator would plan it, with
// one tile shape for the whole batch; the plans are merged into one set of
// device arrays, and the states (padded to at least one tile) sit in one
// HBM arena.  A run is: fill kernels (state |0...0>, queue and pass
// counters), the persistent k_batch launch over every (circuit, pass, tile)
// task, and a per-circuit <Z_q> reduction.  The run can be captured once in
// a HIP graph and replayed.  No counterpart in Kate, which simulates one
// circuit per Simulator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "abi_util.hpp"
#include "kate_hip.h"
#include "kernels.hpp"
#include "plan.hpp"

namespace {

int bfail(int code, const std::string& m) { return kate::set_error(code, m); }

#define KB_HIP(x)                                                                                     \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return bfail(KATE_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int put(T** dst, const std::vector<T>& v)
{
    KB_HIP(hipMalloc((void**)dst, std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) KB_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return KATE_OK;
}

}  // namespace

struct kate_batch {
    int device = 0;
    int dtype = KATE_C64;
    int tile_bits = 12;
    int micro_bits = 4;
    bool wide = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    uint32_t count = 0;
    std::vector<int32_t> nq, nq_int;
    std::vector<int64_t> zoff;
    // final_inverse_permutation of every circuit, concatenated at zoff: the
    // kernel sums <Z> per virtual position p, which is real qubit fip[p]
    std::vector<int32_t> fip;
    int64_t zlen = 0;
    uint64_t total_amps = 0;
    uint32_t ntasks = 0, npasses = 0;
    void* d_states = nullptr;
    uint64_t* d_circ_off = nullptr;
    int32_t* d_nq = nullptr;
    int32_t* d_nq_int = nullptr;
    int64_t* d_zoff = nullptr;
    double* d_z = nullptr;
    double* d_norms = nullptr;
    kate::BatchTask* d_tasks = nullptr;
    uint32_t* d_counters = nullptr;  // done[npasses], next, err
    kate::PassArgs* d_passes = nullptr;
    kate::MicroStep* d_micro = nullptr;
    kate::MicroOp* d_ops = nullptr;
    void* d_coef = nullptr;
    uint32_t* d_raw = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned grid = 0;
    float last_ms = 0;
};

namespace {

void release(kate_batch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void* ptrs[] = {b->d_states, b->d_circ_off, b->d_nq, b->d_nq_int, b->d_zoff, b->d_z, b->d_norms, b->d_tasks,
                    b->d_counters, b->d_passes, b->d_micro, b->d_ops, b->d_coef, b->d_raw};
    for (void* p : ptrs) (void)hipFree(p);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

kate::BatchArgs args(kate_batch* b)
{
    kate::BatchArgs a{};
    a.states = b->d_states;
    a.total_amps = b->total_amps;
    a.circ_off = b->d_circ_off;
    a.nq = b->d_nq;
    a.nq_int = b->d_nq_int;
    a.zoff = b->d_zoff;
    a.z = b->d_z;
    a.norms = b->d_norms;
    a.count = b->count;
    a.tasks = b->d_tasks;
    a.ntasks = b->ntasks;
    a.done = b->d_counters;
    a.npasses = b->npasses;
    a.next = b->d_counters + b->npasses;
    a.err = b->d_counters + b->npasses + 1;
    a.passes = b->d_passes;
    a.dp = kate::DevPlan{b->d_micro, b->d_ops, b->d_coef, b->d_raw};
    a.tile_bits = b->tile_bits;
    a.micro_bits = b->micro_bits;
    a.wide = b->wide;
    a.grid = b->grid;
    return a;
}

}  // namespace

extern "C" {

int kate_batch_create(const kate_program* programs, int32_t count, const kate_options* options, kate_batch** out)
{
    if (!out || !programs || count < 1) return bfail(KATE_ERR_ARG, "bad arguments");
    *out = nullptr;
    kate_options opt{};
    if (options) opt = *options;
    kate::PlanOptions po;
    std::string err = kate::plan_options(opt, po);
    if (!err.empty()) return bfail(KATE_ERR_ARG, err);
    // k_batch runs every pass on the full op set, which has no scaled
    // Euler steps (the light kernel variants have them), with one tile size
    // for every circuit of the batch
    po.factor = false;
    po.auto_tile = false;
    po.scatter = false;
    po.mfma5 = false;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return bfail(KATE_ERR_NO_DEVICE, "no GPU visible");

    kate_batch* b = new kate_batch();
    b->device = opt.device;
    b->dtype = opt.dtype;
    b->tile_bits = po.tile_bits;
    b->micro_bits = po.micro_bits;
    b->count = (uint32_t)count;
    const size_t amp = opt.dtype == KATE_C128 ? 16 : 8;

    std::vector<uint64_t> circ_off;
    std::vector<kate::PassArgs> passes;
    std::vector<kate::MicroStep> micro;
    std::vector<kate::MicroOp> ops;
    std::vector<double> coef;
    std::vector<uint32_t> raw;
    std::vector<uint32_t> pass_base, npass, tiles;
    // per merged pass: the qubits still exactly |0> when it starts (every
    // state starts at |0...0>), and its rest positions
    std::vector<uint64_t> pass_zero;
    for (int32_t c = 0; c < count; c++) {
        kate::Program prog;
        err = kate::program_from_abi(&programs[c], prog);
        if (err.empty() && prog.nqubits > 32) err = "batched circuits have at most 32 qubits";
        kate::Plan plan;
        if (err.empty()) err = kate::build_plan(prog, po, plan);
        if (!err.empty()) {
            delete b;
            return bfail(KATE_ERR_PROGRAM, "circuit " + std::to_string(c) + ": " + err);
        }
        const uint32_t mb = (uint32_t)micro.size(), ob = (uint32_t)ops.size();
        const uint32_t cb = (uint32_t)(coef.size() / 2), rb = (uint32_t)raw.size();
        for (kate::MicroOp o : plan.ops) {
            o.coef += cb;
            ops.push_back(o);
        }
        for (kate::MicroStep m : plan.micro) {
            m.op_begin += ob;
            m.coef += (m.kind == kate::SK_MFMA || m.kind == kate::SK_MFMA_BF16) ? rb : cb;
            if (m.kind != kate::SK_REG) b->wide = true;
            micro.push_back(m);
        }
        coef.insert(coef.end(), plan.coef.begin(), plan.coef.end());
        raw.insert(raw.end(), plan.raw.begin(), plan.raw.end());
        pass_base.push_back((uint32_t)passes.size());
        npass.push_back((uint32_t)plan.passes.size());
        tiles.push_back(1u << (plan.nq_int - plan.tile_bits));
        for (size_t i = 0; i < plan.passes.size(); i++) {
            kate::PassArgs pa = plan.passes[i];
            pa.micro_begin += mb;
            passes.push_back(pa);
            pass_zero.push_back(i < plan.pass_zero.size() && (int)i < plan.support_passes ? plan.pass_zero[i] : 0ull);
        }
        circ_off.push_back(b->total_amps);
        b->total_amps += 1ull << plan.nq_int;
        b->nq.push_back(prog.nqubits);
        b->nq_int.push_back(plan.nq_int);
        b->zoff.push_back(b->zlen);
        b->zlen += prog.nqubits;
        b->fip.insert(b->fip.end(), prog.final_inverse_permutation.begin(), prog.final_inverse_permutation.end());
    }
    b->npasses = (uint32_t)passes.size();
    // tasks by pass level, then circuit, then tile.  Every state starts at
    // |0...0> (written by the fill kernels), so a tile whose rest qubits
    // include one that no gate has touched yet is zero before and after
    // the pass: its task is left out (the memory already holds its zeros).
    std::vector<kate::BatchTask> tasks;
    std::vector<uint32_t> ntask_of_pass(passes.size(), 0);
    const uint32_t maxp = *std::max_element(npass.begin(), npass.end());
    for (uint32_t p = 0; p < maxp; p++)
        for (uint32_t c = 0; c < (uint32_t)count; c++) {
            if (p >= npass[c]) continue;
            const uint32_t gp = pass_base[c] + p;
            const kate::PassArgs& pa = passes[gp];
            uint32_t zero_rest = 0;  // tile-index bits of rest qubits still |0>
            for (int j = 0; j < pa.nrest; j++)
                if ((pass_zero[gp] >> pa.rest_pos[j]) & 1ull) zero_rest |= 1u << j;
            for (uint32_t t = 0; t < tiles[c]; t++) {
                if (t & zero_rest) continue;
                kate::BatchTask k{};
                k.circuit = c;
                k.pass = gp;
                k.tile = t;
                k.wait_pass = p ? gp - 1 : 0xFFFFFFFFu;
                k.wait_count = p ? ntask_of_pass[gp - 1] : 0;
                k.signal = p + 1 < npass[c] ? 1u : 0u;
                tasks.push_back(k);
                ntask_of_pass[gp]++;
            }
        }
    b->ntasks = (uint32_t)tasks.size();

    int rc = KATE_OK;
    auto step = [&](hipError_t e, const char* what) {
        if (rc == KATE_OK && e != hipSuccess)
            rc = bfail(e == hipErrorOutOfMemory ? KATE_ERR_OOM : KATE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    step(hipSetDevice(b->device), "hipSetDevice");
    step(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : b->ev) step(hipEventCreate(&e), "hipEventCreate");
    step(hipMalloc(&b->d_states, b->total_amps * amp), "hipMalloc(states)");
    step(hipMalloc((void**)&b->d_z, std::max<int64_t>(1, b->zlen) * sizeof(double)), "hipMalloc(z)");
    step(hipMalloc((void**)&b->d_norms, b->count * sizeof(double)), "hipMalloc(norms)");
    step(hipMalloc((void**)&b->d_counters, (b->npasses + 2) * sizeof(uint32_t)), "hipMalloc(counters)");
    if (rc == KATE_OK) rc = put(&b->d_circ_off, circ_off);
    if (rc == KATE_OK) rc = put(&b->d_nq, b->nq);
    if (rc == KATE_OK) rc = put(&b->d_nq_int, b->nq_int);
    if (rc == KATE_OK) rc = put(&b->d_zoff, b->zoff);
    if (rc == KATE_OK) rc = put(&b->d_tasks, tasks);
    if (rc == KATE_OK) rc = put(&b->d_passes, passes);
    if (rc == KATE_OK) rc = put(&b->d_micro, micro);
    if (rc == KATE_OK) rc = put(&b->d_ops, ops);
    if (rc == KATE_OK) rc = put(&b->d_raw, raw);
    if (rc == KATE_OK) {
        if (b->dtype == KATE_C128) {
            rc = put((double**)&b->d_coef, coef);
        } else {
            std::vector<float> f(coef.begin(), coef.end());
            rc = put((float**)&b->d_coef, f);
        }
    }
    if (rc == KATE_OK) {
        hipDeviceProp_t prop;
        step(hipGetDeviceProperties(&prop, b->device), "hipGetDeviceProperties");
        // persistent grid: enough workgroups to fill every CU several times over
        const uint64_t per_cu = opt.batch_wgs_per_cu > 0 ? (uint64_t)opt.batch_wgs_per_cu : 8;
        b->grid = (unsigned)std::min<uint64_t>(b->ntasks, (uint64_t)prop.multiProcessorCount * per_cu);
        if (b->grid == 0) b->grid = 1;
    }
    if (rc != KATE_OK) {
        std::string keep = kate_last_error();
        release(b);
        kate::set_error(rc, keep);
        return rc;
    }
    *out = b;
    return KATE_OK;
}

int kate_batch_destroy(kate_batch* b)
{
    release(b);
    return KATE_OK;
}

int kate_batch_capture(kate_batch* b)
{
    if (!b) return bfail(KATE_ERR_ARG, "null batch");
    KB_HIP(hipSetDevice(b->device));
    if (b->exec) return KATE_OK;
    KB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int e = kate::launch_batch(b->dtype, args(b), b->stream);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(b->stream, &g);
    if (e) return bfail(KATE_ERR_HIP, std::string("batch launch in capture: ") + hipGetErrorString((hipError_t)e));
    if (ce != hipSuccess) return bfail(KATE_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    b->graph = g;
    KB_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return KATE_OK;
}

int kate_batch_run(kate_batch* b)
{
    if (!b) return bfail(KATE_ERR_ARG, "null batch");
    KB_HIP(hipSetDevice(b->device));
    KB_HIP(hipEventRecord(b->ev[0], b->stream));
    if (b->exec) {
        KB_HIP(hipGraphLaunch(b->exec, b->stream));
    } else {
        int e = kate::launch_batch(b->dtype, args(b), b->stream);
        if (e) return bfail(KATE_ERR_HIP, std::string("batch launch: ") + hipGetErrorString((hipError_t)e));
    }
    KB_HIP(hipEventRecord(b->ev[1], b->stream));
    KB_HIP(hipEventSynchronize(b->ev[1]));
    KB_HIP(hipEventElapsedTime(&b->last_ms, b->ev[0], b->ev[1]));
    uint32_t errflag = 0;
    KB_HIP(hipMemcpy(&errflag, b->d_counters + b->npasses + 1, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (errflag) return bfail(KATE_ERR_HIP, "batch: a task's wait for its previous pass timed out");
    return KATE_OK;
}

int kate_batch_info(const kate_batch* b, int32_t* count, int64_t* z_len, int32_t* tasks, int32_t* passes,
                    int32_t* grid)
{
    if (!b) return bfail(KATE_ERR_ARG, "null batch");
    if (count) *count = (int32_t)b->count;
    if (z_len) *z_len = b->zlen;
    if (tasks) *tasks = (int32_t)b->ntasks;
    if (passes) *passes = (int32_t)b->npasses;
    if (grid) *grid = (int32_t)b->grid;
    return KATE_OK;
}

int kate_batch_expect_z(kate_batch* b, double* out, int64_t len, double* norms)
{
    if (!b || !out || len != b->zlen) return bfail(KATE_ERR_ARG, "output length must equal the total qubit count");
    KB_HIP(hipSetDevice(b->device));
    std::vector<double> byvirt((size_t)len);
    KB_HIP(hipMemcpy(byvirt.data(), b->d_z, (size_t)len * sizeof(double), hipMemcpyDeviceToHost));
    // virtual position p of circuit c holds real qubit fip[zoff[c] + p]
    for (uint32_t c = 0; c < b->count; c++)
        for (int32_t p = 0; p < b->nq[c]; p++) {
            const size_t at = (size_t)b->zoff[c] + (size_t)p;
            out[(size_t)b->zoff[c] + (size_t)b->fip[at]] = byvirt[at];
        }
    if (norms) KB_HIP(hipMemcpy(norms, b->d_norms, b->count * sizeof(double), hipMemcpyDeviceToHost));
    return KATE_OK;
}

int kate_batch_last_ms(const kate_batch* b, float* ms)
{
    if (!b || !ms) return bfail(KATE_ERR_ARG, "null argument");
    *ms = b->last_ms;
    return KATE_OK;
}

}  // extern "C"
