// kate_ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C-ABI shim around the reference's own CPU build (Kate.hpp compiled with
// g++, the "CPU only" build of the reference README), so that the tests and
// the golden-fixture script can drive Kate's real passes and CPU simulator.
// Built by oracle/Makefile into oracle/_ref/ (never committed) from the
// headers where they lie under the reference checkout; no reference source is
// copied.  Gates carry a tag in Gate::optarg3, a field Kate stores but never
// reads (backendagnosticcode/basic_host_types.hpp:195), so the tests can see
// where each input gate went after scheduling and fusion.
#include "Kate.hpp"
#include "BenchCircuits.hpp"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <iostream>

namespace {
struct Handle {
    Kate::Circuit c;
    explicit Handle(int n) : c(n) {}
};
Handle* H(void* h) { return static_cast<Handle*>(h); }
}  // namespace

extern "C" {

void* kref_new(int nq) { return new Handle(nq); }
void kref_free(void* h) { delete H(h); }

int kref_append(void* h, int id, int nq, const int* qubits, int optarg, const double* mat, double tag)
{
    std::vector<int> q(qubits, qubits + nq);
    if (id == 0) {
        int dim = 1 << nq;
        Kate::Matrix<Kate::Complex> m(dim);
        for (int i = 0; i < dim * dim; i++) m.data[i] = Kate::Complex(mat[2 * i], mat[2 * i + 1]);
        Kate::Gate g(m, q);
        g.optarg = 0;
        g.optarg2 = 0;
        g.optarg3 = Kate::Complex(tag, 0);
        H(h)->c.appendGate(g);
    } else {
        H(h)->c.appendGate(Kate::Gate(id, q, optarg, 0.0, Kate::Complex(tag, 0)));
    }
    return 0;
}

void kref_schedule(void* h) { H(h)->c.gateScheduling(); }
void kref_fuse(void* h, int k, double mtm) { H(h)->c.gateFusion(k, mtm); }
void kref_group(void* h, int g) { H(h)->c.gateGrouping(g); }
void kref_compile_default(void* h, int gpulog2) { H(h)->c.compileDefault(gpulog2); }
void kref_allocate(void* h, int gpulog2) { H(h)->c.allocate(gpulog2); }

int kref_ngates(void* h) { return (int)H(h)->c.gate_set_ordered.size(); }

int kref_gate(void* h, int i, int* id, int* nq, int* qubits, int* optarg, double* mat, double* tag)
{
    Kate::Gate& g = H(h)->c.gate_set_ordered[i];
    *id = g.identifier;
    *nq = (int)g.qbits.size();
    for (int j = 0; j < *nq; j++) qubits[j] = g.qbits[j];
    *optarg = g.identifier == 0 ? 0 : g.optarg;
    *tag = g.optarg3.a;
    if (g.identifier == 0 && mat) {
        int n = g.densecontent.n;
        for (int j = 0; j < n * n; j++) {
            mat[2 * j] = g.densecontent.data[j].a;
            mat[2 * j + 1] = g.densecontent.data[j].b;
        }
    }
    return 0;
}

int kref_ngroups(void* h) { return (int)H(h)->c.groups.size(); }

void kref_groups(void* h, int* ends, unsigned long long* masks)
{
    auto& gr = H(h)->c.groups;
    for (size_t i = 0; i < gr.size(); i++) {
        ends[i] = gr[i].first;
        unsigned long long m = 0;
        for (int q : gr[i].second) m |= 1ull << q;
        masks[i] = m;
    }
}

int kref_ninstr(void* h) { return (int)H(h)->c.instructions.size(); }

int kref_instr(void* h, int i, int* kind, int* pairs)
{
    auto& in = H(h)->c.instructions[i];
    *kind = in.first;
    for (size_t j = 0; j < in.second.size(); j++) pairs[j] = in.second[j];
    return (int)in.second.size();
}

void kref_perms(void* h, int* initial, int* finv)
{
    auto& c = H(h)->c;
    for (size_t i = 0; i < c.initial_permutation.size(); i++) initial[i] = c.initial_permutation[i];
    for (size_t i = 0; i < c.final_inverse_permutation.size(); i++) finv[i] = c.final_inverse_permutation[i];
}

// CPUSimulator::execute on a copy of the circuit (the simulator compiles it
// naively when it has no instructions).  state_out: 2^n complex, virtual order.
int kref_simulate(void* h, const double* proba_in, double* state_out, double* proba_out)
{
    Kate::QuantumCircuit qc = H(h)->c;
    int n = qc.nqbits;
    Kate::CPUSimulator sim(qc, 0);
    Kate::proba_state res(n);
    if (proba_in) {
        Kate::proba_state in(n);
        for (int i = 0; i < n; i++) in.val[i] = std::make_pair(proba_in[2 * i], proba_in[2 * i + 1]);
        res = sim.execute(in);
    } else {
        res = sim.execute();
    }
    if (state_out) std::memcpy(state_out, sim.qbitstate, sizeof(Kate::Complex) * (1ull << n));
    if (proba_out)
        for (int i = 0; i < n; i++) {
            proba_out[2 * i] = res.val[i].first;
            proba_out[2 * i + 1] = res.val[i].second;
        }
    return 0;
}

// CPUSimulator::execute(true) on a copy of the circuit: Kate prints its own
// "Initialization / Computation / measurement time" lines
// (CPUcode/simulator.hpp:86-97); *wall_ms is the wall time of the execute
// call (the simulator's construction and state allocation excluded).
int kref_time_execute(void* h, double* wall_ms)
{
    Kate::QuantumCircuit qc = H(h)->c;
    Kate::CPUSimulator sim(qc, 0);
    const auto t0 = std::chrono::steady_clock::now();
    Kate::proba_state res = sim.execute(true);
    const auto t1 = std::chrono::steady_clock::now();
    std::cout.flush();
    if (wall_ms) *wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return res.val.empty() ? -1 : 0;
}

// Circuit::exportqcx (backendagnosticcode/Circuit.hpp:763-774).
void kref_export_qcx(void* h, const char* path) { H(h)->c.exportqcx(path); }

// CPUmergeGate on gates [first, last) of the circuit.  Returns k.
int kref_merge(void* h, int first, int last, int* out_qubits, double* out_mat)
{
    std::vector<Kate::Gate> v(H(h)->c.gate_set_ordered.begin() + first, H(h)->c.gate_set_ordered.begin() + last);
    Kate::Gate g = Kate::CPUmergeGate(v);
    int k = (int)g.qbits.size();
    for (int j = 0; j < k; j++) out_qubits[j] = g.qbits[j];
    int n = g.densecontent.n;
    for (int j = 0; j < n * n; j++) {
        out_mat[2 * j] = g.densecontent.data[j].a;
        out_mat[2 * j + 1] = g.densecontent.data[j].b;
    }
    return k;
}

// The reference's workload generators (BenchCircuits.hpp:8-58).
void* kref_qft(int n)
{
    Handle* h = new Handle(n);
    h->c = Kate::QFT(n);
    return h;
}

void* kref_qulacs(int n)
{
    Handle* h = new Handle(n);
    h->c = Kate::QulacsBench(n);
    return h;
}

}  // extern "C"
