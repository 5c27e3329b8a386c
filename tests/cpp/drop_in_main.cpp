// A Kate user program, unchanged but for the include: build the circuit with
// Kate's API, run Kate's passes, simulate with Kate's Simulator, print the
// proba_state and the amplitudes (real qubit order).
#include <cstdio>
#include <vector>

#include "Kate.hpp"

using namespace Kate;

int main(int argc, char** argv)
{
    int n = argc > 1 ? atoi(argv[1]) : 10;
    Circuit circuit = QFT(n);
    std::vector<int> q = {0};
    circuit.appendGate(Gate(Hadamard, q));
    q = {1, 0};
    circuit.appendGate(Gate(CNOT, q));
    Matrix<Complex> m(2);
    m(0, 0, Complex(0, 0));
    m(0, 1, Complex(1, 0));
    m(1, 0, Complex(1, 0));
    m(1, 1, Complex(0, 0));
    q = {n - 1};
    circuit.appendGate(Gate(m, q));
    circuit.gateScheduling();
    circuit.gateFusion(3, 0.000001);
    circuit.gateGrouping(8);
    circuit.allocate(0);
    proba_state in(n);
    for (int i = 0; i < n; i++) in.val[i] = std::make_pair((double)(i % 2), 0.0);
    try {
        Simulator sim(circuit, 1);
        proba_state res = sim.execute(in);
        for (int i = 0; i < n; i++) printf("P %d %.17g %.17g\n", i, res.val[i].first, res.val[i].second);
        std::vector<std::complex<double>> st;
        sim.copy_state(st);
        for (size_t i = 0; i < st.size(); i++) printf("A %zu %.17g %.17g\n", i, st[i].real(), st[i].imag());
    } catch (const std::exception& e) {
        printf("ERROR %s\n", e.what());
        return 2;
    }
    return 0;
}
