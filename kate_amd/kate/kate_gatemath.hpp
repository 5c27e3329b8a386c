// kate_gatemath.hpp -- host FP64 gate application and GateMerger fusion.
//
// GateMerger composes a run of gates into one dense gate on the host, in
// FP64 (Circuit::gateFusion always calls the CPU merger, Circuit.hpp:168).
// The composition applies every member gate, in order, to each basis row of an
// identity matrix and transposes the result (CPUcode/GateMerger.hpp:16-54),
// with the per-gate arithmetic of CPUcode/GateComputing.hpp:6-128 (Hadamard as
// a0' = (a0+a1)/sqrt2, a1' = a0' - a1*sqrt2; dense rows summed in column
// order).  Doing the same operations in the same order makes the fused
// matrices bit-identical to Kate's, so fused programs stay comparable.
#ifndef KATE_AMD_GATEMATH_HPP
#define KATE_AMD_GATEMATH_HPP

#include "kate_types.hpp"

namespace Kate {
namespace detail {

// Index of the `line`-th sub-vector base: zero bits inserted at the sorted
// positions `pos[0..k)`.
inline uint64_t spread_zeros(uint64_t line, const int* pos, int k)
{
    for (int i = 0; i < k; i++) {
        uint64_t low = line & ((1ull << pos[i]) - 1ull);
        line = low | ((line - low) << 1);
    }
    return line;
}

// Apply `g` to the 2^nlocal amplitudes `v`, whose bit t is qubit
// local_of[qubit] == t.  Returns false for ids Kate does not compute.
inline bool apply_gate_host(const Gate& g, int nlocal, Complex* v, const std::vector<int>& local_of)
{
    const int k = (int)g.qbits.size();
    if (k < 1 || k > nlocal) return false;
    int lq[64], sorted[64];
    for (int i = 0; i < k; i++) sorted[i] = lq[i] = local_of[(size_t)g.qbits[(size_t)i]];
    std::sort(sorted, sorted + k);
    const uint64_t lines = 1ull << (nlocal - k);
    switch (g.identifier) {
    case TOFFOLI:
        if (k != 3) return false;
        for (uint64_t l = 0; l < lines; l++) {
            uint64_t a = spread_zeros(l, sorted, 3) | (1ull << lq[0]) | (1ull << lq[1]);
            std::swap(v[a], v[a | (1ull << lq[2])]);
        }
        return true;
    case CRk: {
        if (k != 2) return false;
        for (uint64_t l = 0; l < lines; l++) {
            uint64_t a = spread_zeros(l, sorted, 2) | (1ull << lq[0]) | (1ull << lq[1]);
            double ang = ((double)2 * PI) / (double)(1llu << g.optarg);
            v[a] *= Complex(std::cos(ang), std::sin(ang));
        }
        return true;
    }
    case CNOT:
        if (k != 2) return false;
        for (uint64_t l = 0; l < lines; l++) {
            uint64_t a = spread_zeros(l, sorted, 2) | (1ull << lq[1]);
            std::swap(v[a], v[a | (1ull << lq[0])]);
        }
        return true;
    case Hadamard:
        if (k != 1) return false;
        for (uint64_t l = 0; l < lines; l++) {
            uint64_t a = spread_zeros(l, sorted, 1), b = a | (1ull << lq[0]);
            v[a] = (v[a] + v[b]) * SQRT2INV;
            v[b] = v[a] - v[b] * SQRT2;
        }
        return true;
    case 0: {
        const uint64_t dim = 1ull << k;
        if ((uint64_t)g.densecontent.n != dim) return false;
        std::vector<Complex> in(dim);
        std::vector<uint64_t> at(dim);
        for (uint64_t l = 0; l < lines; l++) {
            uint64_t base = spread_zeros(l, sorted, k);
            for (uint64_t s = 0; s < dim; s++) {
                uint64_t x = base;
                for (int i = 0; i < k; i++)
                    if ((s >> i) & 1ull) x |= 1ull << lq[i];
                at[s] = x;
                in[s] = v[x];
            }
            for (uint64_t r = 0; r < dim; r++) {
                Complex acc = 0;
                for (uint64_t c = 0; c < dim; c++) acc += g.densecontent.data[r * dim + c] * in[c];
                v[at[r]] = acc;
            }
        }
        return true;
    }
    default:
        return false;
    }
}

}  // namespace detail

// GateMerger (CPUcode/GateMerger.hpp:8-55): one dense gate on the ascending
// union of the members' qubits.
inline Gate mergeGate(std::vector<Gate> members)
{
    std::set<int> cover;
    int maxq = 0;
    for (const Gate& g : members)
        for (int q : g.qbits) {
            cover.insert(q);
            maxq = std::max(maxq, q);
        }
    std::vector<int> qubits(cover.begin(), cover.end());
    std::vector<int> local_of((size_t)maxq + 1, 0);
    for (size_t i = 0; i < qubits.size(); i++) local_of[(size_t)qubits[i]] = (int)i;
    const int k = (int)qubits.size();
    const int dim = 1 << k;
    Matrix<Complex> m(dim);
    for (int r = 0; r < dim; r++)
        for (int c = 0; c < dim; c++) m(r, c, Complex(r == c ? 1.0 : 0.0));
    for (int r = 0; r < dim; r++)
        for (const Gate& g : members) detail::apply_gate_host(g, k, m.data + (size_t)r * dim, local_of);
    m.transpose();
    return Gate(m, qubits);
}

// The reference's name for the same function in its CPU build (Kate.hpp:14).
inline Gate CPUmergeGate(std::vector<Gate> members) { return mergeGate(std::move(members)); }

}  // namespace Kate

#endif
