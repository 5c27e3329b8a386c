// kate_bench.hpp -- the reference's workload generators
// (BenchCircuits.hpp:8-110): QulacsBench (the published benchmark), QFT
// and randomCircuit, with the same gates in the same order.
#ifndef KATE_AMD_BENCH_HPP
#define KATE_AMD_BENCH_HPP

#include <ctime>

#include "kate_circuit.hpp"

namespace Kate {

// H pairs, nine rounds of H.H.H, ring CNOTs {i, (i+1)%n}: 41 n gates.
inline Circuit QulacsBench(int n)
{
    Circuit c(n);
    std::vector<int> q;
    auto h = [&](int i, int times) {
        for (int t = 0; t < times; t++) {
            q = {i};
            c.appendGate(Gate(Hadamard, q, 0));
        }
    };
    auto ring = [&]() {
        for (int i = 0; i < n; i++) {
            q = {i, (i + 1) % n};
            c.appendGate(Gate(CNOT, q));
        }
    };
    for (int i = 0; i < n; i++) h(i, 2);
    ring();
    for (int a = 0; a < 9; a++) {
        for (int i = 0; i < n; i++) h(i, 3);
        ring();
    }
    for (int i = 0; i < n; i++) h(i, 2);
    return c;
}

// H then CRk(k = j-i+1) on {j, i}; no final swaps, so outputs are bit-reversed.
inline Circuit QFT(int n)
{
    Circuit c(n);
    std::vector<int> q;
    for (int i = 0; i < n; i++) {
        q = {i};
        c.appendGate(Gate(Hadamard, q));
        for (int j = i + 1; j < n; j++) {
            q = {j, i};
            c.appendGate(Gate(CRk, q, j - i + 1));
        }
    }
    return c;
}

// Benchmark-only random circuit: all-ones (non-unitary) dense gates, as in
// the reference; seeded from time() unless a seed is given.
inline Circuit randomCircuit(int n, int depth, unsigned seed = (unsigned)time(NULL))
{
    srand(seed);
    Circuit c(n);
    std::vector<int> q;
    Matrix<Complex> m2(4), m3(8);
    m2.fill(1.);
    m3.fill(1.);
    for (int i = 0; i < depth; i++) {
        int t;
        switch (rand() % 3) {
        case 0:
            q = {rand() % n};
            c.appendGate(Gate(Hadamard, q));
            break;
        case 1:
            t = rand() % n;
            q = {t};
            while (t == q[0]) t = rand() % n;
            q.push_back(t);
            if (rand() % 2 == 1) c.appendGate(Gate(CNOT, q));
            else c.appendGate(Gate(m2, q));
            break;
        default:
            t = rand() % n;
            q = {t};
            while (t == q[0]) t = rand() % n;
            q.push_back(t);
            while (t == q[0] || t == q[1]) t = rand() % n;
            q.push_back(t);
            if (rand() % 2 == 1) c.appendGate(Gate(TOFFOLI, q));
            else c.appendGate(Gate(m3, q));
            break;
        }
    }
    return c;
}

}  // namespace Kate

#endif
