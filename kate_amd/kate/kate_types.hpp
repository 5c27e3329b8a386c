// kate_types.hpp -- Kate's host value types, restated for the MI355X library.
//
// Same names, members and operator semantics as the reference's
// backendagnosticcode/basic_host_types.hpp:5-336 and constants of
// backendagnosticcode/agnosticpreprocessor.hpp:25-33, so that user code
// written against Kate compiles unchanged.  Everything is `inline`, so any
// number of translation units may include it (the reference's non-inline
// definitions allow only one).
#ifndef KATE_AMD_TYPES_HPP
#define KATE_AMD_TYPES_HPP

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <list>
#include <queue>
#include <set>
#include <string>
#include <utility>
#include <vector>

// Public constants and gate ids of Kate's API (macros, as in the reference).
#ifndef PI
#define PI 3.1415926535897932384626433
#endif
#ifndef SQRT2
#define SQRT2 1.4142135623730951
#endif
#ifndef SQRT2INV
#define SQRT2INV 0.7071067811865475
#endif
#ifndef Hadamard
#define Hadamard 2
#endif
#ifndef CNOT
#define CNOT 3
#endif
#ifndef CRk
#define CRk 4
#endif
#ifndef TOFFOLI
#define TOFFOLI 5
#endif
#ifndef RX
#define RX 6
#endif
#ifndef RZ
#define RZ 7
#endif

namespace Kate {

using std::chrono::duration;
using std::chrono::high_resolution_clock;

// Complex number with Kate's rounding: (a+bi)(c+di) = (ac - bd) + (ad + bc)i
// (basic_host_types.hpp:43-45); division and angle as in :54-57 and :31-33.
class Complex {
public:
    double a;
    double b;
    Complex(double re, double im) : a(re), b(im) {}
    Complex(double re) : a(re), b(0) {}
    Complex() : a(0), b(0) {}
    void print() const { std::cout << a << " + " << b << "i"; }
    double norm() const { return std::sqrt(a * a + b * b); }
    // atan, not atan2: the reference's angle() is only defined up to pi.
    double angle() const { return std::atan(b / a); }
    Complex operator+(Complex o) const { return Complex(a + o.a, b + o.b); }
    Complex operator-(Complex o) const { return Complex(a - o.a, b - o.b); }
    Complex operator*(Complex o) const { return Complex(a * o.a - b * o.b, a * o.b + b * o.a); }
    Complex operator*(double o) const { return Complex(a * o, b * o); }
    Complex operator*(int o) const { return Complex(a * o, b * o); }
    Complex operator/(Complex o) const
    {
        double n = o.a * o.a + o.b * o.b;
        return Complex((a * o.a + b * o.b) / n, (o.a * b - a * o.b) / n);
    }
    void operator+=(Complex o)
    {
        a += o.a;
        b += o.b;
    }
    void operator-=(Complex o)
    {
        a -= o.a;
        b -= o.b;
    }
    void operator*=(Complex o)
    {
        double re = a * o.a - b * o.b;
        b = a * o.b + b * o.a;
        a = re;
    }
    void operator*=(double o)
    {
        a *= o;
        b *= o;
    }
};

// Square row-major matrix (basic_host_types.hpp:87-186).  Storage is a
// std::vector, so copies and assignment neither leak nor double-free.
template <typename T>
class Matrix {
public:
    T* data = nullptr;
    int n = 0;

    Matrix() = default;
    explicit Matrix(int dim) : n(dim), store_((size_t)dim * dim) { data = store_.data(); }
    Matrix(const Matrix<T>& o) : n(o.n), store_(o.store_) { data = store_.empty() ? nullptr : store_.data(); }
    Matrix<T>& operator=(const Matrix<T>& o)
    {
        if (this != &o) {
            n = o.n;
            store_ = o.store_;
            data = store_.empty() ? nullptr : store_.data();
        }
        return *this;
    }
    void copy(Matrix<T>& o) { std::copy(o.store_.begin(), o.store_.end(), store_.begin()); }
    void fill(T val) { std::fill(store_.begin(), store_.end(), val); }
    void print()
    {
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n; j++) {
                data[i * n + j].print();
                std::cout << " | ";
            }
            std::cout << std::endl;
        }
    }
    T operator()(int r, int c) const { return data[r * n + c]; }
    void operator()(int r, int c, T val) { data[r * n + c] = val; }
    Matrix<T> operator+(Matrix<T>& o)
    {
        Matrix<T> res(n);
        for (int i = 0; i < n * n; i++) res.data[i] = o.data[i] + data[i];
        return res;
    }
    Matrix<T> operator-(Matrix<T>& o)
    {
        Matrix<T> res(n);
        for (int i = 0; i < n * n; i++) res.data[i] = o.data[i] - data[i];
        return res;
    }
    void operator+=(Matrix<T>& o)
    {
        for (int i = 0; i < n * n; i++) data[i] += o.data[i];
    }
    void operator-=(Matrix<T>& o)
    {
        for (int i = 0; i < n * n; i++) data[i] -= o.data[i];
    }
    Matrix<T> operator*(Matrix<T>& o)
    {
        Matrix<T> res(n);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                T s = 0;
                for (int k = 0; k < n; k++) s += data[i * n + k] * o(k, j);
                res(i, j, s);
            }
        return res;
    }
    void operator*=(double o)
    {
        for (int i = 0; i < n * n; i++) data[i] *= o;
    }
    void transpose()
    {
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) std::swap(data[i * n + j], data[j * n + i]);
    }

private:
    std::vector<T> store_;
};

// A gate (basic_host_types.hpp:190-296).  qbits order is semantic: matrix-
// index bit i acts on qbits[i]; CNOT is {target, control}; TOFFOLI is
// {control, control, target}; CRk's phase is e^{2 pi i / 2^optarg}.
class Gate {
public:
    int identifier = -1;
    int optarg = 0;
    double optarg2 = 0;
    Complex optarg3;
    Matrix<Complex> densecontent;
    std::vector<int> qbits;

    Gate(int id, std::vector<int>& qubits, int oa = 0, double oa2 = 0, Complex oa3 = 0)
        : identifier(id), optarg(oa), optarg2(oa2), optarg3(oa3)
    {
        if (id == 0) std::cout << "you are creating a dense matrix without specifying it. Memory error incoming" << std::endl;
        if (id == Hadamard && qubits.size() != 1) std::cout << "hadamard is on exactly 1 qbit" << std::endl;
        if (id == CNOT && qubits.size() != 2) std::cout << "CNOT is on exactly 2 qbits" << std::endl;
        if (id == CRk && qubits.size() != 2) std::cout << "Controlled Rk is on exactly 2 qbits" << std::endl;
        if (id == TOFFOLI && qubits.size() != 3) std::cout << "toffoli is on exactly 3 qbits" << std::endl;
        if (!unique(qubits)) return;
        qbits = qubits;
    }
    Gate(Matrix<Complex>& dense, std::vector<int>& qubits) : identifier(0), densecontent(dense)
    {
        if (!unique(qubits)) return;
        qbits = qubits;
        if ((1llu << qubits.size()) != (unsigned long long)dense.n)
            std::cout << "size mismatch dense matrix dimension error" << std::endl;
    }
    Gate(int id, Matrix<Complex>& dense, std::vector<int>& qubits) : identifier(id), densecontent(dense)
    {
        if (!unique(qubits)) return;
        qbits = qubits;
        if ((1llu << qubits.size()) != (unsigned long long)dense.n)
            std::cout << "size mismatch dense matrix dimension error" << std::endl;
    }
    void print() const
    {
        std::cout << "Identifier : " << identifier << ", Qbits affected : ";
        for (int q : qbits) std::cout << q << " ";
        std::cout << std::endl;
    }
    // One QCX line (basic_host_types.hpp:270-295).  QCX has no dense gates:
    // they are written as one X per qubit, which loses the matrix.
    std::string transformqcx() const
    {
        std::string s;
        if (identifier == Hadamard) s += "H";
        else if (identifier == 0) {
            std::cout << "Warning: Dense Gate are not supported in QCX, gate will be replaced by XX..X" << std::endl;
            s.append(qbits.size(), 'X');
        } else if (identifier == CNOT) s += "CNOT";
        else if (identifier == CRk) s += "U";
        else if (identifier == TOFFOLI) s += "TOFFOLI";
        s += " ";
        for (int q : qbits) {
            s += std::to_string(q);
            s += " ";
        }
        if (identifier == CRk) s += std::to_string(optarg);
        return s;
    }

private:
    static bool unique(const std::vector<int>& q)
    {
        for (size_t i = 0; i < q.size(); i++)
            for (size_t j = i + 1; j < q.size(); j++)
                if (q[i] == q[j]) {
                    std::cout << "Error while creating a gate: a qbit is present twice: " << q[i] << std::endl;
                    return false;
                }
        return true;
    }
};

inline std::set<int> union_elements(std::set<int>& a, std::set<int>& b)
{
    std::set<int> c = b;
    c.insert(a.begin(), a.end());
    return c;
}

inline std::set<int> union_elements(std::set<int>& a, std::vector<int>& b)
{
    std::set<int> c = a;
    c.insert(b.begin(), b.end());
    return c;
}

inline Complex exp(Complex z)
{
    double m = std::exp(z.a);
    return Complex(m * std::cos(z.b), m * std::sin(z.b));
}

// Per-qubit (theta, phi) pairs (basic_host_types.hpp:319-336).  As input: the
// product state cos(theta pi/2)|0> + e^{i phi} sin(theta pi/2)|1> per qubit.
// As output of execute(): Kate's amplitude-sum summary of the final state.
class proba_state {
public:
    std::vector<std::pair<double, double>> val;
    proba_state(int nqbits) : val((size_t)nqbits, std::make_pair(0.0, 0.0)) {}
    proba_state(std::vector<std::pair<double, double>>& v) : val(v) {}
    void print() const
    {
        for (size_t i = 0; i < val.size(); i++)
            std::cout << "Qbit " << i << " : teta=" << val[i].first << " , phi=" << val[i].second << std::endl;
    }
};

}  // namespace Kate

#endif
