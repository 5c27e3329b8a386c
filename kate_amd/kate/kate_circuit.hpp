// kate_circuit.hpp -- Kate's program format and compiler passes.
//
// QuantumCircuit (backendagnosticcode/QuantumCircuit.hpp:10-205) and Circuit
// (backendagnosticcode/Circuit.hpp:9-354,763-774) with the same members,
// method names and results: gateScheduling, gateFusion (GateMerger),
// gateGrouping, allocate, compileDefault, compileOPT, exportqcx.  Each pass
// reproduces the reference's choices exactly, including its tie-breaking,
// so a program compiled here is the program Kate would run.  The multi-node
// allocators (dual_phase_allocation, slow_fast_allocation, groupallocate,
// Circuit.hpp:355-762) are not part of the single-GPU path and are absent.
#ifndef KATE_AMD_CIRCUIT_HPP
#define KATE_AMD_CIRCUIT_HPP

#include <climits>
#include <fstream>

#include "kate_gatemath.hpp"
#include "kate_graph.hpp"
#include "kate_types.hpp"

namespace Kate {

class QuantumCircuit {
public:
    std::vector<Gate> gate_set_ordered;
    int nqbits = 0;
    // groups[g] = (end-exclusive gate index, qubit set)
    std::vector<std::pair<int, std::set<int>>> groups;
    std::vector<int> initial_permutation;        // real -> virtual
    std::vector<int> final_inverse_permutation;  // virtual -> real
    // (0, flat swap pairs) or (1, {}) = "compute the next group"
    std::vector<std::pair<int, std::vector<int>>> instructions;

    QuantumCircuit(int n) : nqbits(n) {}
    QuantumCircuit(const std::vector<Gate>& gates, int n) : gate_set_ordered(gates), nqbits(n) {}

    void appendGate(const Gate& gate)
    {
        for (int q : gate.qbits)
            if (q >= nqbits) {
                std::cout << "the gate you are trying to add contains a qbit not in the circuit: " << q << "/" << nqbits
                          << "!" << std::endl;
                return;
            }
        if (!groups.empty())
            std::cout << "you should not be adding gate after gateGrouping or allocate, this will not work" << std::endl;
        gate_set_ordered.push_back(gate);
    }

    void print(bool permutationtable = false) const
    {
        if (instructions.empty()) {
            size_t grp = 0;
            if (!groups.empty()) {
                std::cout << "Group 0 with qbits ";
                for (int q : groups[0].second) std::cout << q << " ";
                std::cout << " : " << std::endl;
            }
            for (size_t i = 0; i < gate_set_ordered.size(); i++) {
                if (!groups.empty()) {
                    if (grp < groups.size() && groups[grp].first == (int)i && grp + 1 < groups.size()) {
                        grp++;
                        std::cout << "Group " << grp << " with qbits ";
                        for (int q : groups[grp].second) std::cout << q << " ";
                        std::cout << " : " << std::endl;
                    }
                    std::cout << "   ";
                    gate_set_ordered[i].print();
                } else {
                    std::cout << "gate " << i << " is ";
                    gate_set_ordered[i].print();
                }
            }
            return;
        }
        std::vector<int> inv((size_t)nqbits);
        for (int m = 0; m < nqbits; m++) inv[(size_t)initial_permutation[(size_t)m]] = m;
        std::cout << "-------Quantum-Program-------" << std::endl;
        size_t grp = 0;
        int first = 0;
        for (const auto& ins : instructions) {
            if (ins.first == 0) {
                std::cout << "SWAP ";
                for (size_t m = 0; m + 1 < ins.second.size(); m += 2) {
                    std::cout << ins.second[m] << " and " << ins.second[m + 1] << ", ";
                    std::swap(inv[(size_t)ins.second[m]], inv[(size_t)ins.second[m + 1]]);
                }
            } else if (grp < groups.size()) {
                if (permutationtable) {
                    std::cout << "Permutation table (subjective to real) : " << std::endl;
                    for (int m = 0; m < nqbits; m++) std::cout << m << (m < 10 ? "  " : " ");
                    std::cout << std::endl;
                    for (int m = 0; m < nqbits; m++) std::cout << inv[(size_t)m] << (inv[(size_t)m] < 10 ? "  " : " ");
                    std::cout << std::endl << std::endl;
                }
                std::cout << "EXEC Group " << grp << " with qbits ";
                for (int q : groups[grp].second) std::cout << q << " ";
                std::cout << " : " << std::endl;
                for (int j = first; j < groups[grp].first; j++) {
                    std::cout << "   ";
                    gate_set_ordered[(size_t)j].print();
                }
                first = groups[grp].first;
                grp++;
            }
            std::cout << std::endl << std::endl;
        }
        std::cout << "-----------------------------" << std::endl;
    }

    // Naive compilation (QuantumCircuit.hpp:126-204): one group per gate if
    // ungrouped, identity placement, and a swap with the lowest free local
    // position for every group qubit that sits above max local.
    void compileDefault(int numberofgpulog2 = 0, int maxlocalqbitnumber = 300)
    {
        if (!instructions.empty()) return;
        if (maxlocalqbitnumber + numberofgpulog2 < nqbits) {
            std::cout << "Error: Can't allocate - Too much qbits in the circuit to handle with " << maxlocalqbitnumber
                      << " localqbits and " << (1llu << numberofgpulog2) << " gpus" << std::endl;
            return;
        }
        const int maxlocal = nqbits - numberofgpulog2;
        if (groups.empty()) naive_groups();
        std::vector<int> perm((size_t)nqbits), inv((size_t)nqbits);
        for (int i = 0; i < nqbits; i++) perm[(size_t)i] = inv[(size_t)i] = i;
        initial_permutation = perm;
        int first = 0;
        for (auto& grp : groups) {
            std::vector<int> pairs;
            for (int q : grp.second) {
                if (perm[(size_t)q] < maxlocal) continue;
                int freeq = -1;
                for (int j = 0; j < maxlocal; j++)
                    if (!grp.second.count(inv[(size_t)j])) {
                        freeq = inv[(size_t)j];
                        break;
                    }
                if (freeq < 0) {
                    std::cout << "ALLOCATION FAILED: not enough localqbit available for a given group" << std::endl;
                    return;
                }
                pairs.push_back(perm[(size_t)q]);
                pairs.push_back(perm[(size_t)freeq]);
                std::swap(inv[(size_t)perm[(size_t)q]], inv[(size_t)perm[(size_t)freeq]]);
                std::swap(perm[(size_t)q], perm[(size_t)freeq]);
            }
            if (!pairs.empty()) instructions.push_back(std::make_pair(0, pairs));
            instructions.push_back(std::make_pair(1, std::vector<int>()));
            relabel(grp, first, perm);
            first = grp.first;
        }
        final_inverse_permutation = inv;
    }

protected:
    void naive_groups()
    {
        for (size_t i = 0; i < gate_set_ordered.size(); i++)
            groups.push_back(std::make_pair((int)i + 1, std::set<int>(gate_set_ordered[i].qbits.begin(),
                                                                      gate_set_ordered[i].qbits.end())));
    }
    // Rewrites a group's qubit set and its gates' qubits to virtual positions.
    void relabel(std::pair<int, std::set<int>>& grp, int first, const std::vector<int>& perm)
    {
        std::set<int> moved;
        for (int q : grp.second) moved.insert(perm[(size_t)q]);
        grp.second = moved;
        for (int l = first; l < grp.first; l++)
            for (int& q : gate_set_ordered[(size_t)l].qbits) q = perm[(size_t)q];
    }
};

class Circuit : public QuantumCircuit {
public:
    using QuantumCircuit::QuantumCircuit;

    // Dependency-aware reordering (Circuit.hpp:12-98).  The order is built
    // backwards from the gates nothing depends on; each step picks among the
    // ready gates by the reference's weight |qubits| - |last ∪ qubits|.
    // That weight is computed there in size_t, so it is 0 when the previous
    // gate's qubits are a subset of the candidate's and wraps to ~1.8e19
    // otherwise, and the running best is kept in an int (the wrapped value
    // converts to INT_MIN on x86-64).  The selection rule below reproduces
    // that arithmetic without relying on the out-of-range conversion.
    void gateScheduling()
    {
        const size_t n = gate_set_ordered.size();
        if (n == 0) return;
        std::vector<std::vector<int>> succ(n), pred(n);
        std::vector<int> lastused((size_t)nqbits, -1);
        for (size_t i = 0; i < n; i++)
            for (int q : gate_set_ordered[i].qbits) {
                int& last = lastused[(size_t)q];
                if (last != -1) {
                    std::vector<int>& s = succ[(size_t)last];
                    if (std::find(s.begin(), s.end(), (int)i) == s.end()) {
                        s.push_back((int)i);
                        pred[i].push_back(last);
                    }
                }
                last = (int)i;
            }
        std::list<int> ready;
        std::vector<int> remaining(n);
        for (size_t i = 0; i < n; i++) {
            remaining[i] = (int)succ[i].size();
            if (succ[i].empty()) ready.push_back((int)i);
        }
        std::vector<int> order;
        order.reserve(n);
        std::set<int> last_qbits;
        while (!ready.empty()) {
            auto pick = ready.begin();
            // 0: nothing picked yet (best = -INT32_MAX); 1: best = 0; 2: best = INT_MIN
            int best = 0;
            for (auto it = ready.begin(); it != ready.end(); ++it) {
                const std::vector<int>& q = gate_set_ordered[(size_t)*it].qbits;
                bool subset = true;
                for (int x : last_qbits)
                    if (std::find(q.begin(), q.end(), x) == q.end()) {
                        subset = false;
                        break;
                    }
                // weight 0 beats -INT32_MAX and INT_MIN; the wrapped weight beats everything but itself... and 0.
                bool better = subset ? (best != 1) : true;
                if (better) {
                    pick = it;
                    best = subset ? 1 : 2;
                }
            }
            int u = *pick;
            ready.erase(pick);
            order.push_back(u);
            last_qbits = std::set<int>(gate_set_ordered[(size_t)u].qbits.begin(), gate_set_ordered[(size_t)u].qbits.end());
            for (int p : pred[(size_t)u])
                if (--remaining[(size_t)p] == 0) ready.push_back(p);
        }
        std::vector<Gate> reordered;
        reordered.reserve(order.size());
        for (size_t i = order.size(); i-- > 0;) reordered.push_back(gate_set_ordered[(size_t)order[i]]);
        gate_set_ordered = reordered;
    }

    // GateMerger fusion (Circuit.hpp:99-175): shortest path over gate
    // prefixes.  Keeping gate i alone costs 2^k (dense), 2 (H), 0.75 (CNOT),
    // 0.25 (CRk) or 3/8 (TOFFOLI); merging gates i..j costs 2^|union| plus
    // merge_time_matters * sum_{m=i+1..j} 2^|q_m| * 4^|union| / 2^n, for
    // unions of at most qbitsizelimit qubits.  Each merged run becomes one
    // dense gate composed in FP64 by mergeGate.
    void gateFusion(int qbitsizelimit = 5, double merge_time_matters = 0.000001)
    {
        const int n = (int)gate_set_ordered.size();
        SparsePath graph(n + 1);
        for (int i = 0; i < n; i++) {
            const Gate& g = gate_set_ordered[(size_t)i];
            std::set<int> cover(g.qbits.begin(), g.qbits.end());
            switch (g.identifier) {
            case 0: graph.edge(i, i + 1, (double)(1llu << cover.size())); break;
            case Hadamard: graph.edge(i, i + 1, 2); break;
            case CNOT: graph.edge(i, i + 1, 0.75); break;
            case CRk: graph.edge(i, i + 1, 0.25); break;
            case TOFFOLI: graph.edge(i, i + 1, (double)3 / 8); break;
            default: break;
            }
            double members = 0;
            for (int j = i + 1; j < n; j++) {
                cover.insert(gate_set_ordered[(size_t)j].qbits.begin(), gate_set_ordered[(size_t)j].qbits.end());
                if ((int)cover.size() > qbitsizelimit) break;
                members += (double)(1llu << gate_set_ordered[(size_t)j].qbits.size());
                double cost = (double)(1llu << cover.size());
                cost += merge_time_matters * members * (double)(1llu << (2 * cover.size())) / (double)(1llu << nqbits);
                graph.edge(i, j + 1, cost);
            }
        }
        std::vector<int> path = graph.path(0, n);
        std::vector<Gate> fused;
        int from = -1;
        for (int node : path) {
            if (from == -1) {
                from = node;
                continue;
            }
            if (node == from + 1) {
                fused.push_back(gate_set_ordered[(size_t)from]);
            } else {
                std::vector<Gate> run(gate_set_ordered.begin() + from, gate_set_ordered.begin() + node);
                fused.push_back(mergeGate(run));
            }
            from = node;
        }
        gate_set_ordered = fused;
    }

    // Fewest contiguous gate ranges whose qubit union stays within
    // qbitgroupsize (Circuit.hpp:176-206), by unit-weight shortest path.
    void gateGrouping(int qbitgroupsize = 8)
    {
        const int n = (int)gate_set_ordered.size();
        if (n == 0) return;
        SparsePath graph(n + 1);
        for (int i = 0; i < n; i++) {
            std::set<int> cover;
            for (int j = i; j < n; j++) {
                cover.insert(gate_set_ordered[(size_t)j].qbits.begin(), gate_set_ordered[(size_t)j].qbits.end());
                if ((int)cover.size() > qbitgroupsize) break;
                graph.edge(i, j + 1, 1);
            }
        }
        std::vector<int> cuts = graph.path(0, n);
        if (cuts.size() < 2) {
            std::cout << "Error: a gate is larger than the group size " << qbitgroupsize << std::endl;
            return;
        }
        groups.assign(cuts.size() - 1, std::make_pair(0, std::set<int>()));
        size_t grp = 1;
        std::set<int> cover;
        for (int i = 0; i < n; i++) {
            if (grp < cuts.size() && cuts[grp] == i) {
                groups[grp - 1] = std::make_pair(i, cover);
                cover.clear();
                grp++;
            }
            cover.insert(gate_set_ordered[(size_t)i].qbits.begin(), gate_set_ordered[(size_t)i].qbits.end());
        }
        groups.back() = std::make_pair(n, cover);
    }

    // Qubit placement (Circuit.hpp:207-348): qubits first used earliest are
    // local; a group needing a non-local qubit swaps it with the local qubit
    // whose next use is furthest away (Belady).  Qubits are relabelled even
    // with no GPUs, so the state comes out in virtual order.
    void allocate(int numberofgpulog2 = 0, int maxlocalqbitnumber = 300)
    {
        if (maxlocalqbitnumber + numberofgpulog2 < nqbits) {
            std::cout << "Error: Can't allocate - Too much qbits in the circuit to handle with " << maxlocalqbitnumber
                      << " localqbits and " << (1llu << numberofgpulog2) << " gpus" << std::endl;
            return;
        }
        const int maxlocal = nqbits - numberofgpulog2;
        instructions.clear();
        if (groups.empty()) naive_groups();
        const int G = (int)groups.size();
        // next_use[i] = (qubit, distance to its next group after i)
        std::vector<std::set<std::pair<int, int>>> next_use((size_t)G);
        std::vector<int> last_seen((size_t)nqbits, INT32_MAX);
        for (int i = G - 1; i >= 0; i--)
            for (int q : groups[(size_t)i].second) {
                next_use[(size_t)i].insert(std::make_pair(q, last_seen[(size_t)q] - i));
                last_seen[(size_t)q] = i;
            }
        std::vector<int> by_first_use((size_t)nqbits);
        for (int i = 0; i < nqbits; i++) by_first_use[(size_t)i] = i;
        std::sort(by_first_use.begin(), by_first_use.end(),
                  [&last_seen](int a, int b) { return last_seen[(size_t)a] < last_seen[(size_t)b]; });
        std::vector<int> nextsee(last_seen);
        std::vector<int> perm((size_t)nqbits, 0), inv((size_t)nqbits, 0);
        for (int i = 0; i < nqbits; i++) {
            perm[(size_t)by_first_use[(size_t)i]] = i;
            inv[(size_t)i] = by_first_use[(size_t)i];
        }
        initial_permutation = perm;
        int first = 0;
        for (int i = 0; i < G; i++) {
            for (int& x : nextsee) x -= 1;
            std::vector<int> pairs;
            const std::set<int> taken = groups[(size_t)i].second;
            for (int q : groups[(size_t)i].second) {
                if (perm[(size_t)q] >= maxlocal) {
                    int victim = -1;
                    for (int j = 0; j < maxlocal; j++) {
                        int c = inv[(size_t)j];
                        if (taken.count(c)) continue;
                        if (victim == -1 || nextsee[(size_t)c] > nextsee[(size_t)victim]) victim = c;
                    }
                    if (victim == -1) {
                        std::cout << "ALLOCATION FAILED: not enough localqbit available for a given group" << std::endl;
                        return;
                    }
                    pairs.push_back(perm[(size_t)q]);
                    pairs.push_back(perm[(size_t)victim]);
                    std::swap(inv[(size_t)perm[(size_t)q]], inv[(size_t)perm[(size_t)victim]]);
                    std::swap(perm[(size_t)q], perm[(size_t)victim]);
                }
                nextsee[(size_t)q] = INT32_MAX;
            }
            for (const auto& nu : next_use[(size_t)i]) nextsee[(size_t)nu.first] = nu.second;
            if (!pairs.empty()) instructions.push_back(std::make_pair(0, pairs));
            instructions.push_back(std::make_pair(1, std::vector<int>()));
            relabel(groups[(size_t)i], first, perm);
            first = groups[(size_t)i].first;
        }
        final_inverse_permutation = inv;
    }

    void compileOPT(int qbit_matrix_merge_size_limit = 5, double merge_time_matters = 0.000001, int groupsize = 10,
                    int numberofgpulog2 = 0, int maxlocalqbitnumber = 300)
    {
        gateScheduling();
        gateFusion(qbit_matrix_merge_size_limit, merge_time_matters);
        gateGrouping(groupsize);
        allocate(numberofgpulog2, maxlocalqbitnumber);
    }

    // QCX text export (Circuit.hpp:763-774).
    void exportqcx(std::string filename)
    {
        std::ofstream file(filename);
        if (!file) {
            std::cout << "Error: Could not open " << filename << " to write qcx file" << std::endl;
            return;
        }
        file << "QUBITS " << nqbits << std::endl;
        for (const Gate& g : gate_set_ordered) file << g.transformqcx() << std::endl;
        file << "BEGIN MEASUREMENT";
    }
};

}  // namespace Kate

#endif
