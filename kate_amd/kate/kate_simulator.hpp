// kate_simulator.hpp -- Kate's Simulator on one MI355X.
//
// Same constructor and execute() overloads as the reference's GPU simulator
// (HIPcode/simulator.hpp:220-543): Simulator(QuantumCircuit, number_of_gpu,
// swapBufferSizeLog2) compiles an uncompiled circuit with compileDefault,
// allocates the state and uploads the program once; execute() initialises,
// runs every instruction and returns Kate's proba_state.  Underneath it lowers
// the circuit onto the C ABI of libkate_hip.so (include/kate_hip.h).
//
// Differences, on purpose:
//  * no silent CPU fallback: without a usable GPU the constructor throws
//    std::runtime_error (the reference prints a warning and runs its CPU
//    simulator, HIPcode/simulator.hpp:245-255);
//  * programs compiled for several GPUs run on one: their swap instructions
//    become in-HBM qubit remaps;
//  * execute(proba_state&) honours its input (the reference's CPU fallback
//    drops it, HIPcode/simulator.hpp:355);
//  * copy_state / expect_z / norm / timings are additions.
#ifndef KATE_AMD_SIMULATOR_HPP
#define KATE_AMD_SIMULATOR_HPP

#include <complex>
#include <stdexcept>

#include "kate_circuit.hpp"
#include "kate_hip.h"

namespace Kate {

// A QuantumCircuit flattened into the arrays a kate_program points at.
struct LoweredProgram {
    std::vector<kate_gate_rec> gates;
    std::vector<double> pool;
    std::vector<int32_t> group_end;
    std::vector<uint64_t> group_mask;
    std::vector<int32_t> instr_kind, instr_npairs, instr_off, pairs;
    std::vector<int32_t> initial, final_inverse;
    kate_program prog{};

    explicit LoweredProgram(const QuantumCircuit& qc)
    {
        for (const Gate& g : qc.gate_set_ordered) {
            kate_gate_rec r{};
            r.id = g.identifier;
            r.nq = (int32_t)g.qbits.size();
            for (size_t i = 0; i < g.qbits.size() && i < KATE_MAX_GATE_QUBITS; i++) r.qubits[i] = g.qbits[i];
            r.optarg = g.optarg;
            r.matrix_offset = -1;
            if (g.identifier == 0) {
                r.matrix_offset = (int64_t)(pool.size() / 2);
                const int n = g.densecontent.n;
                for (int i = 0; i < n * n; i++) {
                    pool.push_back(g.densecontent.data[i].a);
                    pool.push_back(g.densecontent.data[i].b);
                }
            }
            gates.push_back(r);
        }
        for (const auto& grp : qc.groups) {
            group_end.push_back(grp.first);
            uint64_t m = 0;
            for (int q : grp.second) m |= 1ull << q;
            group_mask.push_back(m);
        }
        for (const auto& ins : qc.instructions) {
            instr_kind.push_back(ins.first);
            instr_off.push_back((int32_t)pairs.size());
            instr_npairs.push_back((int32_t)(ins.second.size() / 2));
            pairs.insert(pairs.end(), ins.second.begin(), ins.second.end());
        }
        initial.assign(qc.initial_permutation.begin(), qc.initial_permutation.end());
        final_inverse.assign(qc.final_inverse_permutation.begin(), qc.final_inverse_permutation.end());
        prog.nqubits = qc.nqbits;
        prog.ngates = (int32_t)gates.size();
        prog.gates = gates.data();
        prog.matrix_pool = pool.data();
        prog.matrix_pool_len = (int64_t)(pool.size() / 2);
        prog.ngroups = (int32_t)group_end.size();
        prog.group_end = group_end.data();
        prog.group_mask = group_mask.data();
        prog.ninstr = (int32_t)instr_kind.size();
        prog.instr_kind = instr_kind.data();
        prog.instr_npairs = instr_npairs.data();
        prog.instr_pair_offset = instr_off.data();
        prog.pairs = pairs.data();
        prog.initial_permutation = initial.empty() ? nullptr : initial.data();
        prog.final_inverse_permutation = final_inverse.empty() ? nullptr : final_inverse.data();
    }
    LoweredProgram(const LoweredProgram&) = delete;
    LoweredProgram& operator=(const LoweredProgram&) = delete;
};

// printGpuInfo (HIPcode/DeviceInfo.hpp:10-36).
inline void printGpuInfo()
{
    char buf[4096];
    if (kate_device_info(buf, sizeof(buf)) == KATE_OK) std::cout << buf;
    else std::cout << "no GPU: " << kate_last_error() << std::endl;
}

class Simulator {
public:
    std::vector<std::pair<int, std::set<int>>> groups;
    std::vector<int> initial_permutation;
    std::vector<int> final_inverse_permutation;
    std::vector<std::pair<int, std::vector<int>>> instructions;
    std::vector<Gate> gate_set_ordered;
    int nqbits = 0;
    int number_of_gpu = 1;
    Complex** gpu_qbits_states = nullptr;  // gpu_qbits_states[0]: device address of the state
    int swapBufferSizeLog2 = 24;           // kept for API parity; remaps need no staging buffer

    Simulator(QuantumCircuit mycircuit, int number_of_gpu_, int swapBufferSizeLog2_ = 24)
        : Simulator(std::move(mycircuit), number_of_gpu_, swapBufferSizeLog2_, default_options())
    {
    }

    Simulator(QuantumCircuit mycircuit, int number_of_gpu_, int swapBufferSizeLog2_, const kate_options& opts)
        : swapBufferSizeLog2(swapBufferSizeLog2_), dtype_(opts.dtype)
    {
        int count = 0;
        if (kate_device_count(&count) != KATE_OK || count == 0)
            throw std::runtime_error(std::string("Kate::Simulator: no usable GPU: ") + kate_last_error());
        if (number_of_gpu_ > 1)
            std::cout << "Warning: this build simulates on one GPU; swap instructions run as in-HBM qubit remaps"
                      << std::endl;
        if (mycircuit.instructions.empty()) {
            std::cout << "warning: the simulator has been input a circuit that is not compiled. I will compile it naively now"
                      << std::endl;
            int g = 0;
            while ((1 << (g + 1)) <= number_of_gpu_) g++;
            mycircuit.compileDefault(g, mycircuit.nqbits - g);
        }
        groups = mycircuit.groups;
        initial_permutation = mycircuit.initial_permutation;
        final_inverse_permutation = mycircuit.final_inverse_permutation;
        instructions = mycircuit.instructions;
        gate_set_ordered = mycircuit.gate_set_ordered;
        nqbits = mycircuit.nqbits;
        LoweredProgram lp(mycircuit);
        int rc = kate_sim_create(&lp.prog, &opts, &sim_);
        if (rc != KATE_OK) throw std::runtime_error(std::string("Kate::Simulator: ") + kate_last_error());
        void* d = nullptr;
        kate_sim_device_state(sim_, &d);
        state_ptr_[0] = static_cast<Complex*>(d);
        gpu_qbits_states = state_ptr_;
    }

    ~Simulator()
    {
        if (sim_) kate_sim_destroy(sim_);
    }
    Simulator(const Simulator&) = delete;
    Simulator& operator=(const Simulator&) = delete;

    proba_state execute(bool displaytime = false) { return run(nullptr, displaytime); }

    proba_state execute(proba_state& in, bool displaytime = false)
    {
        if ((int)in.val.size() != nqbits) {
            std::cout << "wrong input proba_state_size_input, defaulting to no input" << std::endl;
            return run(nullptr, displaytime);
        }
        std::vector<double> tp;
        for (const auto& v : in.val) {
            tp.push_back(v.first);
            tp.push_back(v.second);
        }
        return run(tp.data(), displaytime);
    }

    // Amplitudes (complex128 regardless of the state's precision), real qubit
    // order unless real_order is false.
    void copy_state(std::vector<std::complex<double>>& out, bool real_order = true)
    {
        out.resize(1ull << nqbits);
        if (dtype_ == KATE_C128) {
            check(kate_sim_read_state(sim_, out.data(), out.size() * sizeof(std::complex<double>), real_order ? 1 : 0));
        } else {
            std::vector<std::complex<float>> f(out.size());
            check(kate_sim_read_state(sim_, f.data(), f.size() * sizeof(std::complex<float>), real_order ? 1 : 0));
            for (size_t i = 0; i < f.size(); i++) out[i] = std::complex<double>(f[i].real(), f[i].imag());
        }
    }

    std::vector<double> expect_z(const std::vector<uint64_t>& masks)
    {
        std::vector<double> out(masks.size());
        check(kate_sim_expect_z(sim_, masks.data(), (int32_t)masks.size(), out.data()));
        return out;
    }

    double norm()
    {
        double n = 0;
        check(kate_sim_norm(sim_, &n));
        return n;
    }

    kate_timings timings()
    {
        kate_timings t{};
        check(kate_sim_last_timings(sim_, &t));
        return t;
    }

    kate_sim* handle() { return sim_; }

    // Kate's callers read the amplitudes through gpu_qbits_states right after
    // execute(), so the default options run every pass inside execute()
    // (no deferred trailing permutation passes); a caller that only reads
    // through this class or the C ABI may pass options without
    // KATE_PLAN_NO_DEFER.
    static kate_options default_options()
    {
        kate_options o{};
        o.dtype = KATE_C128;
        o.plan_flags = KATE_PLAN_NO_DEFER;
        return o;
    }

private:
    kate_sim* sim_ = nullptr;
    Complex* state_ptr_[1] = {nullptr};
    int dtype_ = KATE_C128;

    static void check(int rc)
    {
        if (rc != KATE_OK) throw std::runtime_error(std::string("Kate::Simulator: ") + kate_last_error());
    }

    proba_state run(const double* in, bool displaytime)
    {
        std::vector<double> tp(2 * (size_t)nqbits);
        check(kate_sim_execute(sim_, in, tp.data()));
        // a plan that ends with an out-of-place pass leaves the state in the
        // simulator's other buffer: gpu_qbits_states[0] follows it
        void* d = nullptr;
        check(kate_sim_device_state(sim_, &d));
        state_ptr_[0] = static_cast<Complex*>(d);
        std::vector<std::pair<double, double>> v((size_t)nqbits);
        for (int i = 0; i < nqbits; i++) v[(size_t)i] = std::make_pair(tp[2 * (size_t)i], tp[2 * (size_t)i + 1]);
        if (displaytime) {
            kate_timings t = timings();
            std::cout << "Initialization/Measurement : " << t.init_ms << " / " << t.measure_ms << " ms" << std::endl;
            std::cout << "Total computing time : " << t.compute_ms << " ms in " << t.passes << " passes ("
                      << t.remap_passes << " remaps)" << std::endl;
        }
        return proba_state(v);
    }
};

}  // namespace Kate

#endif
