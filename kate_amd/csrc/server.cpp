// server.cpp -- continuous-batching request loop (kate_server_* C ABI).
//
// No counterpart in Kate, whose Simulator runs one circuit to completion per
// call.  A server owns one HBM arena (a single hipMalloc at creation), a pool
// of HIP streams and one scheduler thread:
//   * submit() copies the program, Z-string masks and optional product-state
//     input and queues the request (first come, first served);
//   * the scheduler plans queued requests, admits the oldest ones whose
//     footprint (state + plan + reduction scratch) fits the free arena,
//     enqueues init, every pass, the Z reduction and the result copy on the
//     next stream of the pool, and keeps going: requests that arrive while
//     others run are admitted as soon as memory frees up, and their kernels
//     overlap on the device through the different streams;
//   * completed requests (event query) release their arena ranges and wake
//     their waiters.
// Every request runs the same kernels it would run alone, so its result does
// not depend on what else is in flight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "abi_util.hpp"
#include "kate_hip.h"
#include "kernels.hpp"
#include "plan.hpp"

namespace {

int sfail(int code, const std::string& m) { return kate::set_error(code, m); }

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, size) in 256-byte units; host bookkeeping only.
class Arena {
public:
    explicit Arena(uint64_t size) { free_[0] = size; }
    bool alloc(uint64_t bytes, uint64_t& off)
    {
        bytes = (bytes + 255) & ~255ull;
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes) {
                off = it->first;
                const uint64_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                return true;
            }
        return false;
    }
    void release(uint64_t off, uint64_t bytes)
    {
        bytes = (bytes + 255) & ~255ull;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t largest() const
    {
        uint64_t m = 0;
        for (const auto& f : free_) m = std::max(m, f.second);
        return m;
    }

private:
    std::map<uint64_t, uint64_t> free_;
};

struct Request {
    uint64_t id = 0;
    kate::Program prog;
    kate::Plan plan;
    std::vector<uint64_t> masks;  // virtual-position masks, [0] = 0 (norm)
    std::vector<double> theta_phi;  // product-state input (2 per qubit) or empty
    std::vector<double> product;    // its factors (complex128) over the plan's qubits, set when planned
    std::string error;
    bool planned = false;
    // arena ranges
    uint64_t off = 0, bytes = 0;
    uint64_t o_state = 0, o_micro = 0, o_ops = 0, o_coef = 0, o_raw = 0, o_scratch = 0, o_out = 0, o_masks = 0,
             o_factors = 0, o_state2 = 0, o_scatter = 0;
    std::vector<float> coef32;
    int stream = -1;
    hipEvent_t done_ev = nullptr;
    double* host_out = nullptr;  // pinned slot
    int slot = -1;
    // outcome
    std::vector<double> z;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
};

// Kate's product-state input (theta, phi per qubit; simulator.hpp's
// initialize(proba_state)) as one (|0>, |1>) factor pair per state qubit of
// the plan, qubits the program does not have left at |0>.
static void product_factors(Request& r)
{
    const int n = r.prog.nqubits, m = r.plan.nq_int;
    r.product.assign(4 * (size_t)m, 0.0);
    for (int j = 0; j < m; j++) r.product[2 * (size_t)j] = 1.0;
    const double pi = 3.1415926535897932384626433;
    for (int i = 0; i < n; i++) {
        const double th = r.theta_phi[2 * (size_t)i], ph = r.theta_phi[2 * (size_t)i + 1];
        const int v = r.prog.initial_permutation[(size_t)i];
        const double sn = std::sin(th * pi / 2);
        r.product[2 * (size_t)v] = std::cos(th * pi / 2);
        r.product[2 * (size_t)v + 1] = 0.0;
        r.product[2 * ((size_t)v + m)] = std::cos(ph) * sn;
        r.product[2 * ((size_t)v + m) + 1] = std::sin(ph) * sn;
    }
}

}  // namespace

struct kate_server {
    kate_server_options opt{};
    size_t amp = 16;
    char* arena_base = nullptr;
    std::unique_ptr<Arena> arena;
    std::vector<hipStream_t> streams;
    int next_stream = 0;
    double* pinned = nullptr;  // max_active slots of 64 doubles
    std::vector<int> free_slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::vector<std::shared_ptr<Request>> active;
    std::map<uint64_t, std::shared_ptr<Request>> all;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    kate::PlanOptions po;
    uint64_t completed = 0;
    std::string fatal;
};

namespace {

// Requests of this many qubits or more get a second state buffer and
// out-of-place (relabelled) passes; smaller ones keep one buffer.
constexpr int kServeOutOfPlaceQubits = 22;

uint64_t round256(uint64_t x) { return (x + 255) & ~255ull; }

// Footprint layout of a planned request inside its arena range.
uint64_t layout(kate_server* s, Request& r)
{
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t at = o;
        o += round256(std::max<uint64_t>(bytes, 1));
        return at;
    };
    r.o_state = take((1ull << r.plan.nq_int) * s->amp);
    r.o_micro = take(r.plan.micro.size() * sizeof(kate::MicroStep));
    r.o_ops = take(r.plan.ops.size() * sizeof(kate::MicroOp));
    r.o_coef = take(r.plan.coef.size() * (s->amp / 2));
    r.o_raw = take(r.plan.raw.size() * sizeof(uint32_t));
    r.o_scratch = take(kate::reduce_scratch_doubles(r.plan.nq_int) * sizeof(double));
    r.o_out = take(64 * sizeof(double));
    r.o_masks = take(16 * sizeof(uint64_t));
    r.o_factors = take(4 * (uint64_t)r.plan.nq_int * sizeof(double));
    if (!r.plan.scatter_args.empty()) {  // out-of-place passes: a second state buffer
        r.o_state2 = take((1ull << r.plan.nq_int) * s->amp);
        r.o_scatter = take(r.plan.scatter_args.size() * sizeof(kate::ScatterArgs));
    }
    return o;
}

// Enqueues everything a request does on its stream; returns a HIP error code.
int launch_request(kate_server* s, Request& r)
{
    hipStream_t st = s->streams[(size_t)r.stream];
    char* base = s->arena_base + r.off;
    void* state = base + r.o_state;
    hipError_t e;
#define SV(x)                  \
    if ((e = (x)) != hipSuccess) \
        return (int)e;
    SV(hipMemcpyAsync(base + r.o_micro, r.plan.micro.data(), r.plan.micro.size() * sizeof(kate::MicroStep),
                      hipMemcpyHostToDevice, st));
    SV(hipMemcpyAsync(base + r.o_ops, r.plan.ops.data(), r.plan.ops.size() * sizeof(kate::MicroOp),
                      hipMemcpyHostToDevice, st));
    if (s->opt.dtype == KATE_C128) {
        SV(hipMemcpyAsync(base + r.o_coef, r.plan.coef.data(), r.plan.coef.size() * sizeof(double),
                          hipMemcpyHostToDevice, st));
    } else {
        SV(hipMemcpyAsync(base + r.o_coef, r.coef32.data(), r.coef32.size() * sizeof(float), hipMemcpyHostToDevice,
                          st));
    }
    if (!r.plan.raw.empty())
        SV(hipMemcpyAsync(base + r.o_raw, r.plan.raw.data(), r.plan.raw.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice, st));
    SV(hipMemcpyAsync(base + r.o_masks, r.masks.data(), r.masks.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    if (!r.plan.scatter_args.empty())
        SV(hipMemcpyAsync(base + r.o_scatter, r.plan.scatter_args.data(),
                          r.plan.scatter_args.size() * sizeof(kate::ScatterArgs), hipMemcpyHostToDevice, st));
    int k = 0;
    if (!r.product.empty()) {
        SV(hipMemcpyAsync(base + r.o_factors, r.product.data(), r.product.size() * sizeof(double),
                          hipMemcpyHostToDevice, st));
        k = kate::launch_init_product(s->opt.dtype, state, r.plan.nq_int, (const double*)(base + r.o_factors), st);
    } else if (r.plan.passes.empty()) {
        k = kate::launch_init_basis(s->opt.dtype, state, r.plan.nq_int, 0, st);
    }
    if (k) return k;
    // from |0...0> the first pass makes the state and the passes follow its
    // support (no init kernel); a product state is read as written
    const bool from_zero = r.product.empty();
    kate::DevPlan dp{(const kate::MicroStep*)(base + r.o_micro), (const kate::MicroOp*)(base + r.o_ops),
                     base + r.o_coef, (const uint32_t*)(base + r.o_raw)};
    // an out-of-place pass (relabelled, or storing through the trailing
    // permutation) writes the other buffer, which then holds the state
    void* other = r.plan.scatter_args.empty() ? nullptr : base + r.o_state2;
    const kate::ScatterArgs* sa = (const kate::ScatterArgs*)(base + r.o_scatter);
    for (size_t i = 0; i < r.plan.passes.size(); i++) {
        const kate::PassArgs pa = from_zero ? kate::zero_start_pass(r.plan, i, s->po.support, r.plan.passes.size()) : r.plan.passes[i];
        if (!pa.scatter) {
            k = kate::launch_pass(s->opt.dtype, r.plan.tile_bits, state, r.plan.nq_int, pa, dp, st);
        } else if (!other) {
            k = (int)hipErrorInvalidValue;
        } else if (pa.micro_count) {
            kate::DevPlan d = dp;
            d.dst = other;
            d.scatter = sa;
            k = kate::launch_pass(s->opt.dtype, r.plan.tile_bits, state, r.plan.nq_int, pa, d, st);
        } else {
            k = kate::launch_scatter(s->opt.dtype, r.plan.tile_bits, state, other, pa, sa, st);
        }
        if (k) return k;
        if (pa.scatter) std::swap(state, other);
    }
    k = kate::launch_zsums(s->opt.dtype, state, r.plan.nq_int, (const uint64_t*)(base + r.o_masks),
                           (int)r.masks.size(), (double*)(base + r.o_scratch), (double*)(base + r.o_out), st);
    if (k) return k;
    SV(hipMemcpyAsync(r.host_out, base + r.o_out, 16 * sizeof(double), hipMemcpyDeviceToHost, st));
    SV(hipEventRecord(r.done_ev, st));
#undef SV
    return 0;
}

void finish(kate_server* s, Request& r, bool ok)
{
    if (ok) {
        r.z.resize(r.masks.size() - 1);
        const double norm = r.host_out[0];
        for (size_t j = 1; j < r.masks.size(); j++) r.z[j - 1] = r.host_out[j] / norm;
        r.state = 2;
    } else {
        r.state = 3;
    }
    r.t_done = now_s();
    s->arena->release(r.off, r.bytes);
    s->free_slots.push_back(r.slot);
    (void)hipEventDestroy(r.done_ev);
    r.done_ev = nullptr;
    s->completed++;
}

void worker_loop(kate_server* s)
{
    (void)hipSetDevice(s->opt.device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        // plan queued requests with the lock released (submit() may run meanwhile)
        std::vector<std::shared_ptr<Request>> unplanned;
        for (auto& r : s->queue)
            if (!r->planned) unplanned.push_back(r);
        if (!unplanned.empty()) {
            lk.unlock();
            for (auto& r : unplanned) {
                // requests of kServeOutOfPlaceQubits or more run out of
                // place (two state buffers): their passes are HBM-bound and
                // relabelled passes read whole tile blocks
                kate::PlanOptions po = s->po;
                po.scatter = r->prog.nqubits >= kServeOutOfPlaceQubits;
                // a product-state input has no support region to follow:
                // its passes can all be relabelled
                po.zero_start = r->theta_phi.empty();
                std::string err = kate::build_plan(r->prog, po, r->plan);
                if (err.empty() && s->opt.dtype == KATE_C64) r->coef32.assign(r->plan.coef.begin(), r->plan.coef.end());
                if (err.empty() && !r->theta_phi.empty()) product_factors(*r);
                if (!err.empty()) r->error = "planning: " + err;
            }
            lk.lock();
            for (auto& r : unplanned) r->planned = true;
        }
        // admit, oldest first, while memory and slots allow
        while (!s->queue.empty() && (int)s->active.size() < s->opt.max_active && !s->free_slots.empty()) {
            auto r = s->queue.front();
            if (!r->planned) break;
            if (!r->error.empty()) {
                s->queue.pop_front();
                r->state = 3;
                r->t_done = now_s();
                s->cv_done.notify_all();
                continue;
            }
            r->bytes = layout(s, *r);
            if (!s->arena->alloc(r->bytes, r->off)) {
                if (s->active.empty()) {  // can never fit
                    s->queue.pop_front();
                    r->error = "request does not fit the server's HBM budget";
                    r->state = 3;
                    r->t_done = now_s();
                    s->cv_done.notify_all();
                    continue;
                }
                break;
            }
            s->queue.pop_front();
            r->slot = s->free_slots.back();
            s->free_slots.pop_back();
            r->host_out = s->pinned + 64 * r->slot;
            r->stream = s->next_stream;
            s->next_stream = (s->next_stream + 1) % (int)s->streams.size();
            r->t_admit = now_s();
            r->state = 1;
            (void)hipEventCreateWithFlags(&r->done_ev, hipEventDisableTiming);
            int e = launch_request(s, *r);
            if (e) {
                r->error = std::string("launch: ") + hipGetErrorString((hipError_t)e);
                (void)hipStreamSynchronize(s->streams[(size_t)r->stream]);
                finish(s, *r, false);
                s->cv_done.notify_all();
                continue;
            }
            s->active.push_back(r);
        }
        // retire completed requests
        bool progressed = false;
        for (size_t i = 0; i < s->active.size();) {
            auto& r = s->active[i];
            hipError_t q = hipEventQuery(r->done_ev);
            if (q == hipErrorNotReady) {
                i++;
                continue;
            }
            if (q != hipSuccess) r->error = std::string("device: ") + hipGetErrorString(q);
            finish(s, *r, q == hipSuccess);
            s->active.erase(s->active.begin() + (long)i);
            progressed = true;
        }
        if (progressed) s->cv_done.notify_all();
        if (s->stop && s->queue.empty() && s->active.empty()) break;
        if (s->active.empty() && s->queue.empty())
            s->cv_work.wait(lk);
        else if (!progressed)
            s->cv_work.wait_for(lk, std::chrono::microseconds(100));
    }
}

void release_server(kate_server* s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (hipStream_t st : s->streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    (void)hipFree(s->arena_base);
    (void)hipHostFree(s->pinned);
    delete s;
}

}  // namespace

extern "C" {

int kate_server_create(const kate_server_options* options, kate_server** out)
{
    if (!out) return sfail(KATE_ERR_ARG, "null output");
    *out = nullptr;
    kate_server_options o{};
    if (options) o = *options;
    if (o.streams <= 0) o.streams = 4;
    if (o.streams > 16) o.streams = 16;
    if (o.max_active <= 0) o.max_active = 32;
    if (o.max_active > 256) o.max_active = 256;
    kate_options ko{};
    ko.dtype = o.dtype;
    ko.tile_qubits = o.tile_qubits;
    ko.low_qubits = o.low_qubits;
    ko.use_mfma = o.use_mfma;
    ko.micro_bits = o.micro_bits;
    ko.device = o.device;
    kate::PlanOptions po;
    std::string err = kate::plan_options(ko, po);
    if (!err.empty()) return sfail(KATE_ERR_ARG, err);
    po.scatter = false;  // set per request, by size
    if (po.auto_tile) {
        // one tile size, planned once per request on the scheduler thread:
        // 12-qubit tiles (a mix of 128 circuits of 12..30 qubits at
        // saturation: 202 circuits/s against 181 with 13-qubit tiles and 143
        // with both planned per request)
        po.auto_tile = false;
        po.tile_bits = 12;
        po.min_qubits = 12;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return sfail(KATE_ERR_NO_DEVICE, "no GPU visible");
    if (o.hbm_budget_bytes <= 0) {
        size_t fr = 0, tot = 0;
        (void)hipSetDevice(o.device);
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return sfail(KATE_ERR_HIP, "hipMemGetInfo failed");
        o.hbm_budget_bytes = (int64_t)(fr / 2);
    }
    kate_server* s = new kate_server();
    s->opt = o;
    s->po = po;
    s->amp = o.dtype == KATE_C128 ? 16 : 8;
    hipError_t e = hipSetDevice(o.device);
    if (e == hipSuccess) e = hipMalloc((void**)&s->arena_base, (size_t)o.hbm_budget_bytes);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->pinned, (size_t)o.max_active * 64 * sizeof(double), 0);
    for (int i = 0; e == hipSuccess && i < o.streams; i++) {
        hipStream_t st;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) s->streams.push_back(st);
    }
    if (e != hipSuccess) {
        const std::string m = std::string("server setup: ") + hipGetErrorString(e);
        release_server(s);
        return sfail(e == hipErrorOutOfMemory ? KATE_ERR_OOM : KATE_ERR_HIP, m);
    }
    s->arena.reset(new Arena((uint64_t)o.hbm_budget_bytes));
    for (int i = o.max_active - 1; i >= 0; i--) s->free_slots.push_back(i);
    s->worker = std::thread(worker_loop, s);
    *out = s;
    return KATE_OK;
}

int kate_server_destroy(kate_server* s)
{
    release_server(s);
    return KATE_OK;
}

int kate_server_submit(kate_server* s, const kate_program* program, const double* theta_phi, const uint64_t* zmasks,
                       int32_t nmasks, uint64_t* id)
{
    if (!s || !program || !id || nmasks < 0 || nmasks > 15 || (nmasks && !zmasks))
        return sfail(KATE_ERR_ARG, "bad arguments (at most 15 Z-string masks per request)");
    auto r = std::make_shared<Request>();
    std::string err = kate::program_from_abi(program, r->prog);
    if (!err.empty()) return sfail(KATE_ERR_PROGRAM, err);
    const int n = r->prog.nqubits;
    r->masks.push_back(0);
    for (int j = 0; j < nmasks; j++) {
        if (n < 64 && (zmasks[j] >> n)) return sfail(KATE_ERR_ARG, "mask names a qubit outside the circuit");
        uint64_t v = 0;
        for (int p = 0; p < n; p++)
            if ((zmasks[j] >> r->prog.final_inverse_permutation[(size_t)p]) & 1ull) v |= 1ull << p;
        r->masks.push_back(v);
    }
    if (theta_phi) r->theta_phi.assign(theta_phi, theta_phi + 2 * (size_t)n);
    {
        std::lock_guard<std::mutex> g(s->mu);
        r->id = s->next_id++;
        r->t_submit = now_s();
        s->queue.push_back(r);
        s->all[r->id] = r;
        *id = r->id;
    }
    s->cv_work.notify_all();
    return KATE_OK;
}

int kate_server_wait(kate_server* s, uint64_t id, double* zout, int32_t cap, kate_request_stats* st)
{
    if (!s) return sfail(KATE_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return sfail(KATE_ERR_ARG, "unknown request id");
    auto r = it->second;
    s->cv_done.wait(lk, [&] { return r->state >= 2; });
    if (st) {
        st->submit_s = r->t_submit;
        st->admit_s = r->t_admit;
        st->done_s = r->t_done;
        st->latency_ms = (r->t_done - r->t_submit) * 1e3;
        st->passes = (int32_t)r->plan.passes.size();
        st->nqubits = r->prog.nqubits;
        st->footprint_bytes = (int64_t)r->bytes;
    }
    s->all.erase(it);
    if (r->state == 3) return sfail(KATE_ERR_HIP, r->error.empty() ? "request failed" : r->error);
    for (int j = 0; j < cap && j < (int)r->z.size(); j++) zout[j] = r->z[(size_t)j];
    return KATE_OK;
}

int kate_server_pending(kate_server* s, int32_t* queued, int32_t* active, int64_t* completed, int64_t* free_bytes)
{
    if (!s) return sfail(KATE_ERR_ARG, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    if (queued) *queued = (int32_t)s->queue.size();
    if (active) *active = (int32_t)s->active.size();
    if (completed) *completed = (int64_t)s->completed;
    if (free_bytes) *free_bytes = (int64_t)s->arena->largest();
    return KATE_OK;
}

}  // extern "C"
