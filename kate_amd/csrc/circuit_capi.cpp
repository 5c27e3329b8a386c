// circuit_capi.cpp -- the kate_circuit_* C ABI: Kate's circuit compiler
// (backendagnosticcode/Circuit.hpp, QuantumCircuit.hpp) as implemented by the
// header-only layer kate_amd/kate/, exported for C and ctypes callers.
#include <cstring>
#include <memory>
#include <string>

#include "Kate.hpp"
#include "kate_hip.h"

struct kate_circuit {
    Kate::Circuit c;
    explicit kate_circuit(int n) : c(n) {}
};

namespace kate {
int set_error(int code, const std::string& msg);
}

namespace {
int cfail(int code, const std::string& m) { return kate::set_error(code, m); }
}  // namespace

extern "C" {

int kate_circuit_create(int32_t nqubits, kate_circuit** out)
{
    if (!out || nqubits < 1 || nqubits > KATE_MAX_QUBITS) return cfail(KATE_ERR_ARG, "bad arguments");
    *out = new kate_circuit(nqubits);
    return KATE_OK;
}

int kate_circuit_destroy(kate_circuit* c)
{
    delete c;
    return KATE_OK;
}

int kate_circuit_append(kate_circuit* c, int32_t id, int32_t nq, const int32_t* qubits, int32_t optarg,
                        const double* matrix)
{
    if (!c || nq < 1 || nq > KATE_MAX_GATE_QUBITS || !qubits) return cfail(KATE_ERR_ARG, "bad gate arguments");
    std::vector<int> q(qubits, qubits + nq);
    uint64_t seen = 0;
    for (int x : q) {
        if (x < 0 || x >= c->c.nqbits) return cfail(KATE_ERR_ARG, "qubit out of range");
        if (seen & (1ull << x)) return cfail(KATE_ERR_ARG, "a qubit is present twice");
        seen |= 1ull << x;
    }
    if (id != 0 && id != Hadamard && id != CNOT && id != CRk && id != TOFFOLI)
        return cfail(KATE_ERR_ARG, "gate id must be 0, 2, 3, 4 or 5");
    if (id == 0) {
        if (!matrix || nq > 13) return cfail(KATE_ERR_ARG, "dense gate needs a matrix of at most 13 qubits");
        const int dim = 1 << nq;
        Kate::Matrix<Kate::Complex> m(dim);
        for (int i = 0; i < dim * dim; i++) m.data[i] = Kate::Complex(matrix[2 * i], matrix[2 * i + 1]);
        c->c.appendGate(Kate::Gate(m, q));
    } else {
        c->c.appendGate(Kate::Gate(id, q, optarg));
    }
    return KATE_OK;
}

int kate_circuit_schedule(kate_circuit* c)
{
    if (!c) return KATE_ERR_ARG;
    c->c.gateScheduling();
    return KATE_OK;
}

int kate_circuit_fuse(kate_circuit* c, int32_t max_qubits, double mtm)
{
    if (!c || max_qubits < 1 || max_qubits > 13) return KATE_ERR_ARG;
    c->c.gateFusion(max_qubits, mtm);
    return KATE_OK;
}

int kate_circuit_group(kate_circuit* c, int32_t g)
{
    if (!c || g < 1) return KATE_ERR_ARG;
    c->c.gateGrouping(g);
    return KATE_OK;
}

int kate_circuit_allocate(kate_circuit* c, int32_t gpu_log2, int32_t max_local)
{
    if (!c || gpu_log2 < 0) return KATE_ERR_ARG;
    c->c.allocate(gpu_log2, max_local);
    return KATE_OK;
}

int kate_circuit_compile_default(kate_circuit* c, int32_t gpu_log2, int32_t max_local)
{
    if (!c || gpu_log2 < 0) return KATE_ERR_ARG;
    c->c.compileDefault(gpu_log2, max_local);
    return KATE_OK;
}

int kate_circuit_compile_opt(kate_circuit* c, int32_t fuse_qubits, double mtm, int32_t group_qubits, int32_t gpu_log2,
                             int32_t max_local)
{
    if (!c) return KATE_ERR_ARG;
    c->c.compileOPT(fuse_qubits, mtm, group_qubits, gpu_log2, max_local);
    return KATE_OK;
}

int kate_circuit_export_qcx(kate_circuit* c, const char* path)
{
    if (!c || !path) return KATE_ERR_ARG;
    c->c.exportqcx(path);
    return KATE_OK;
}

int kate_circuit_info(const kate_circuit* c, int32_t* nq, int32_t* ng, int32_t* ngr, int32_t* ni)
{
    if (!c) return KATE_ERR_ARG;
    if (nq) *nq = c->c.nqbits;
    if (ng) *ng = (int32_t)c->c.gate_set_ordered.size();
    if (ngr) *ngr = (int32_t)c->c.groups.size();
    if (ni) *ni = (int32_t)c->c.instructions.size();
    return KATE_OK;
}

int kate_circuit_gate(const kate_circuit* c, int32_t i, int32_t* id, int32_t* nq, int32_t* qubits, int32_t* optarg,
                      double* matrix)
{
    if (!c || i < 0 || i >= (int32_t)c->c.gate_set_ordered.size()) return KATE_ERR_ARG;
    const Kate::Gate& g = c->c.gate_set_ordered[(size_t)i];
    if (id) *id = g.identifier;
    if (nq) *nq = (int32_t)g.qbits.size();
    if (qubits)
        for (size_t j = 0; j < g.qbits.size(); j++) qubits[j] = g.qbits[j];
    if (optarg) *optarg = g.identifier == 0 ? 0 : g.optarg;
    if (matrix && g.identifier == 0) {
        const int n = g.densecontent.n;
        for (int j = 0; j < n * n; j++) {
            matrix[2 * j] = g.densecontent.data[j].a;
            matrix[2 * j + 1] = g.densecontent.data[j].b;
        }
    }
    return KATE_OK;
}

int kate_circuit_groups(const kate_circuit* c, int32_t* ends, uint64_t* masks)
{
    if (!c) return KATE_ERR_ARG;
    for (size_t i = 0; i < c->c.groups.size(); i++) {
        if (ends) ends[i] = c->c.groups[i].first;
        uint64_t m = 0;
        for (int q : c->c.groups[i].second) m |= 1ull << q;
        if (masks) masks[i] = m;
    }
    return KATE_OK;
}

int kate_circuit_instruction(const kate_circuit* c, int32_t i, int32_t* kind, int32_t* pairs, int32_t cap,
                             int32_t* npairs_ints)
{
    if (!c || i < 0 || i >= (int32_t)c->c.instructions.size()) return KATE_ERR_ARG;
    const auto& ins = c->c.instructions[(size_t)i];
    if (kind) *kind = ins.first;
    if (npairs_ints) *npairs_ints = (int32_t)ins.second.size();
    for (size_t j = 0; j < ins.second.size() && (int32_t)j < cap; j++)
        if (pairs) pairs[j] = ins.second[j];
    return KATE_OK;
}

int kate_circuit_permutations(const kate_circuit* c, int32_t* initial, int32_t* final_inverse)
{
    if (!c) return KATE_ERR_ARG;
    for (size_t i = 0; i < c->c.initial_permutation.size(); i++)
        if (initial) initial[i] = c->c.initial_permutation[i];
    for (size_t i = 0; i < c->c.final_inverse_permutation.size(); i++)
        if (final_inverse) final_inverse[i] = c->c.final_inverse_permutation[i];
    return KATE_OK;
}

int kate_sim_create_from_circuit(const kate_circuit* c, const kate_options* options, kate_sim** out)
{
    if (!c || !out) return KATE_ERR_ARG;
    Kate::QuantumCircuit qc = c->c;
    if (qc.instructions.empty()) qc.compileDefault(0, qc.nqbits);
    Kate::LoweredProgram lp(qc);
    return kate_sim_create(&lp.prog, options, out);
}

int kate_circuit_plan(const kate_circuit* c, const kate_options* options, int32_t* passes, int32_t* remap_passes,
                      char* buf, size_t len)
{
    if (!c) return KATE_ERR_ARG;
    Kate::QuantumCircuit qc = c->c;
    if (qc.instructions.empty()) qc.compileDefault(0, qc.nqbits);
    Kate::LoweredProgram lp(qc);
    return kate_plan_describe(&lp.prog, options, passes, remap_passes, buf, len);
}

int kate_batch_create_from_circuits(const kate_circuit* const* circuits, int32_t count, const kate_options* options,
                                    kate_batch** out)
{
    if (!circuits || count < 1 || !out) return cfail(KATE_ERR_ARG, "bad arguments");
    std::vector<std::unique_ptr<Kate::LoweredProgram>> lowered;
    std::vector<kate_program> progs;
    for (int32_t i = 0; i < count; i++) {
        if (!circuits[i]) return cfail(KATE_ERR_ARG, "null circuit");
        Kate::QuantumCircuit qc = circuits[i]->c;
        if (qc.instructions.empty()) qc.compileDefault(0, qc.nqbits);
        lowered.emplace_back(new Kate::LoweredProgram(qc));
        progs.push_back(lowered.back()->prog);
    }
    return kate_batch_create(progs.data(), count, options, out);
}

int kate_server_submit_circuit(kate_server* server, const kate_circuit* c, const double* theta_phi,
                               const uint64_t* zmasks, int32_t nmasks, uint64_t* id)
{
    if (!c) return cfail(KATE_ERR_ARG, "null circuit");
    Kate::QuantumCircuit qc = c->c;
    if (qc.instructions.empty()) qc.compileDefault(0, qc.nqbits);
    Kate::LoweredProgram lp(qc);
    return kate_server_submit(server, &lp.prog, theta_phi, zmasks, nmasks, id);
}

}  // extern "C"
