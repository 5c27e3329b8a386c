// plan.cpp -- lowers a Kate program onto passes over the state.
//
// Kate runs one kernel per compute group, each a full read+write of the state
// (HIPcode/simulator.hpp:760-810), with the groups chosen on the host by
// gateGrouping.  Here the groups only delimit swap instructions: the gates
// between two swaps are re-packed into as few passes as the tile allows.
//
// A pass stages 2^tile_bits amplitudes per workgroup in LDS.  Its tile
// always contains qubits 0..low-1, so every HBM access is a contiguous run of
// 2^low amplitudes, plus the qubits its gates need.  Gates may leave program
// order only past gates they commute with (disjoint qubits): a gate runs in a
// pass when all earlier gates sharing a qubit with it ran before.  Inside a
// pass the gates are packed the same way into register micro-steps of 4 tile
// bits (16 amplitudes per thread, one LDS round trip per micro-step).
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <sstream>

namespace kate {
namespace {

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

// Which gates of `masks` (in order) run when only the qubits in `set` are
// available: a gate runs if its qubits are in `set` and no earlier gate
// sharing a qubit with it was left out.  Gates with ok[i] == 0 never run.
int sweep(const std::vector<uint64_t>& masks, const std::vector<uint8_t>& ok, uint64_t set, uint64_t universe,
          std::vector<int>* frontier, std::vector<int>* accepted)
{
    uint64_t blocked = 0;
    int count = 0;
    for (size_t i = 0; i < masks.size(); i++) {
        const uint64_t m = masks[i];
        if (m & blocked) {
            blocked |= m;
        } else if (ok[i] && (m & ~set) == 0) {
            count++;
            if (accepted) accepted->push_back((int)i);
        } else {
            if (frontier && ok[i]) frontier->push_back((int)i);
            blocked |= m;
        }
        if ((blocked & universe) == universe) break;
    }
    return count;
}

// The sweep of register micro-steps.  CNOTs (lazy[i]) never need their
// qubits in registers: one whose qubits are held and not pending runs as a
// register op, any other is applied to the tile layout after the step; a
// gate after a layout CNOT on one of its qubits waits for the next step.
int sweep_micro(const std::vector<uint64_t>& masks, const std::vector<uint8_t>& ok, const std::vector<uint8_t>& lazy,
                uint64_t set, std::vector<int>* frontier, std::vector<int>* reg, std::vector<int>* late)
{
    uint64_t blocked = 0, pending = 0;
    int count = 0;
    for (size_t i = 0; i < masks.size(); i++) {
        const uint64_t m = masks[i];
        if (m & blocked) {
            blocked |= m;
        } else if (lazy[i]) {
            count++;
            if ((m & ~set) == 0 && !(m & pending)) {
                if (reg) reg->push_back((int)i);
            } else {
                pending |= m;
                if (late) late->push_back((int)i);
            }
        } else if (ok[i] && (m & ~set) == 0 && !(m & pending)) {
            count++;
            if (reg) reg->push_back((int)i);
        } else {
            if (frontier && ok[i] && !(m & pending)) frontier->push_back((int)i);
            blocked |= m;
        }
    }
    return count;
}

uint64_t choose_micro_set(const std::vector<uint64_t>& masks, const std::vector<uint8_t>& ok,
                          const std::vector<uint8_t>& lazy, int width)
{
    uint64_t set = 0;
    for (;;) {
        std::vector<int> frontier;
        const int now = sweep_micro(masks, ok, lazy, set, &frontier, nullptr, nullptr);
        double best = 0;
        uint64_t best_new = 0;
        std::vector<uint64_t> tried;
        for (int f : frontier) {
            const uint64_t add = masks[(size_t)f] & ~set;
            if (popc(set | add) > width) continue;
            if (std::find(tried.begin(), tried.end(), add) != tried.end()) continue;
            tried.push_back(add);
            const int gain = sweep_micro(masks, ok, lazy, set | add, nullptr, nullptr, nullptr) - now;
            const double score = (double)gain / popc(add);
            if (score > best) {
                best = score;
                best_new = add;
            }
        }
        if (!best_new) return set;
        set |= best_new;
    }
}

// Greedy growth of a qubit set up to `capacity`: repeatedly add the new
// qubits of the ready gate that makes the most gates runnable per added
// qubit (earliest gate on ties).
uint64_t choose_set(const std::vector<uint64_t>& masks, const std::vector<uint8_t>& ok, int capacity, uint64_t start,
                    uint64_t universe)
{
    uint64_t set = start;
    for (;;) {
        std::vector<int> frontier;
        const int now = sweep(masks, ok, set, universe, &frontier, nullptr);
        double best = 0;
        uint64_t best_new = 0;
        std::vector<uint64_t> tried;
        for (int f : frontier) {
            const uint64_t add = masks[(size_t)f] & ~set;
            if (popc(set | add) > capacity) continue;
            if (std::find(tried.begin(), tried.end(), add) != tried.end()) continue;
            tried.push_back(add);
            const int gain = sweep(masks, ok, set | add, universe, nullptr, nullptr) - now;
            const double score = (double)gain / popc(add);
            if (score > best) {
                best = score;
                best_new = add;
            }
        }
        if (!best_new) return set;
        set |= best_new;
    }
}

// Lowest positions below `limit` not in `set`, added until popc == want.
uint64_t pad_set(uint64_t set, int want, int limit)
{
    for (int b = 0; b < limit && popc(set) < want; b++) set |= 1ull << b;
    return set;
}

// M'[r'][c'] = M[r][c], where new matrix bit j is old bit order[j].
std::vector<double> permute_matrix(const double* m, int k, const int* order)
{
    const uint64_t dim = 1ull << k;
    std::vector<double> out(2 * dim * dim);
    auto remap = [&](uint64_t x) {
        uint64_t y = 0;
        for (int j = 0; j < k; j++)
            if ((x >> j) & 1ull) y |= 1ull << order[j];
        return y;
    };
    for (uint64_t r = 0; r < dim; r++)
        for (uint64_t c = 0; c < dim; c++) {
            const uint64_t src = remap(r) * dim + remap(c);
            out[2 * (r * dim + c)] = m[2 * src];
            out[2 * (r * dim + c) + 1] = m[2 * src + 1];
        }
    return out;
}

// Splits dense 2-qubit gates that are a Kronecker product of two 1-qubit
// gates, optionally times one CNOT on either side (P * (A x B) or
// (A x B) * P, P a CNOT between the two qubits or the identity), into those
// factors.  Kate's gateFusion produces such gates when it merges the 1-qubit
// gates of two neighbouring qubits with the CNOT between them
// (backendagnosticcode/Circuit.hpp:99-175); as one dense 4x4 they need both
// qubits in the same register step and tie every gate before and after them
// on either qubit together, while the factors are composed with their
// neighbours on each qubit and the CNOT costs nothing (it becomes a layout
// update).  A gate is split only when the factors reproduce it to 1e-15
// max-abs (FP64 composition rounding), so results change by rounding only.
bool kron_factors(const double* m, double (&a)[8], double (&b)[8])
{
    // R[(r1,c1)][(r0,c0)] = M[r1*2+r0][c1*2+c0] (index bit 0 = qubits[0])
    double R[4][4][2];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            const int i = ((r >> 1) << 1) | (c >> 1), j = ((r & 1) << 1) | (c & 1);
            R[i][j][0] = m[2 * (r * 4 + c)];
            R[i][j][1] = m[2 * (r * 4 + c) + 1];
        }
    int bi = 0, bj = 0;
    double best = -1;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            const double v = R[i][j][0] * R[i][j][0] + R[i][j][1] * R[i][j][1];
            if (v > best) {
                best = v;
                bi = i;
                bj = j;
            }
        }
    if (best <= 0) return false;
    // a = column bj of R, b = row bi of R / R[bi][bj]
    const double pr = R[bi][bj][0], pi = R[bi][bj][1], pn = pr * pr + pi * pi;
    double av[4][2], bv[4][2];
    for (int i = 0; i < 4; i++) {
        av[i][0] = R[i][bj][0];
        av[i][1] = R[i][bj][1];
        bv[i][0] = (R[bi][i][0] * pr + R[bi][i][1] * pi) / pn;
        bv[i][1] = (R[bi][i][1] * pr - R[bi][i][0] * pi) / pn;
    }
    double na = 0, nb = 0;
    for (int i = 0; i < 4; i++) {
        na += av[i][0] * av[i][0] + av[i][1] * av[i][1];
        nb += bv[i][0] * bv[i][0] + bv[i][1] * bv[i][1];
    }
    const double t = std::sqrt(std::sqrt(nb / na));  // balance the two factors
    for (int i = 0; i < 4; i++) {
        a[2 * i] = av[i][0] * t;
        a[2 * i + 1] = av[i][1] * t;
        b[2 * i] = bv[i][0] / t;
        b[2 * i + 1] = bv[i][1] / t;
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            const double re = a[2 * i] * b[2 * j] - a[2 * i + 1] * b[2 * j + 1];
            const double im = a[2 * i] * b[2 * j + 1] + a[2 * i + 1] * b[2 * j];
            if (std::fabs(re - R[i][j][0]) > 2e-15 || std::fabs(im - R[i][j][1]) > 2e-15) return false;
        }
    return true;
}

// The 4x4 permutation of a CNOT between the two qubits of a 2-qubit gate
// (matrix index bit i = qubits[i]); ctl = the control's index bit.
int cnot_perm(int ctl, int r)
{
    const int tgt = 1 - ctl;
    return ((r >> ctl) & 1) ? r ^ (1 << tgt) : r;
}

// Program with every splittable dense 2-qubit gate replaced by its factors
// (group ends follow the gate indices).
Program split_kron_gates(const Program& in)
{
    Program out = in;
    out.gates.clear();
    std::vector<int> new_index(in.gates.size() + 1, 0);
    for (size_t gi = 0; gi < in.gates.size(); gi++) {
        new_index[gi] = (int)out.gates.size();
        const PlanGate& g = in.gates[gi];
        bool done = false;
        if (g.id == 0 && g.k == 2) {
            const double* m = in.mat.data() + 2 * g.moff;
            // forms: M = K (ctl -1), M = P K (P after the factors), M = K P (P first)
            for (int form = 0; form < 5 && !done; form++) {
                const int ctl = form == 0 ? -1 : (form - 1) % 2;
                const bool p_first = form >= 3;
                double k4[32];
                for (int r = 0; r < 4; r++)
                    for (int c = 0; c < 4; c++) {
                        // K = P M (P after) or K = M P (P first); P is an involution
                        const int rr = (ctl >= 0 && !p_first) ? cnot_perm(ctl, r) : r;
                        const int cc = (ctl >= 0 && p_first) ? cnot_perm(ctl, c) : c;
                        k4[2 * (r * 4 + c)] = m[2 * (rr * 4 + cc)];
                        k4[2 * (r * 4 + c) + 1] = m[2 * (rr * 4 + cc) + 1];
                    }
                double a[8], b[8];
                if (!kron_factors(k4, a, b)) continue;
                // a acts on index bit 1 (qubits[1]), b on bit 0 (qubits[0])
                PlanGate cx{};
                cx.id = 3;
                cx.k = 2;
                if (ctl >= 0) {
                    cx.q[0] = g.q[1 - ctl];  // Kate's CNOT is {target, control}
                    cx.q[1] = g.q[ctl];
                    cx.mask = (1ull << cx.q[0]) | (1ull << cx.q[1]);
                }
                auto one = [&](const double* u, int q) {
                    PlanGate h{};
                    h.id = 0;
                    h.k = 1;
                    h.q[0] = q;
                    h.mask = 1ull << q;
                    h.moff = (int64_t)(out.mat.size() / 2);
                    out.mat.insert(out.mat.end(), u, u + 8);
                    out.gates.push_back(h);
                };
                if (ctl >= 0 && p_first) out.gates.push_back(cx);
                one(b, g.q[0]);
                one(a, g.q[1]);
                if (ctl >= 0 && !p_first) out.gates.push_back(cx);
                done = true;
            }
        }
        if (!done) out.gates.push_back(g);
    }
    new_index[in.gates.size()] = (int)out.gates.size();
    for (int& e : out.group_end) e = new_index[(size_t)e];
    return out;
}

struct Builder {
    const Program& prog;
    const PlanOptions& opt;
    Plan& plan;
    uint64_t zero = ~0ull;  // qubits still exactly |0> in a run from |0...0>
    bool remapped = false;  // a remap pass was planned: support no longer tracked

    // Support update for one gate (sound: a qubit stays in `zero` only if
    // every amplitude with that bit set is still exactly zero).
    void track(const PlanGate& g)
    {
        switch (g.id) {
        case 3:  // CNOT {target, control}: moves weight only if the control can be 1
            if (!((zero >> g.q[1]) & 1ull)) zero &= ~(1ull << g.q[0]);
            break;
        case 5:  // TOFFOLI {c0, c1, target}
            if (!((zero >> g.q[0]) & 1ull) && !((zero >> g.q[1]) & 1ull)) zero &= ~(1ull << g.q[2]);
            break;
        case 4:  // CRk: diagonal
            break;
        case 0: {
            const uint64_t dim = 1ull << g.k;
            const double* m = prog.mat.data() + 2 * g.moff;
            bool diag = true;
            for (uint64_t r = 0; r < dim && diag; r++)
                for (uint64_t c = 0; c < dim; c++)
                    if (r != c && (m[2 * (r * dim + c)] != 0.0 || m[2 * (r * dim + c) + 1] != 0.0)) {
                        diag = false;
                        break;
                    }
            if (!diag) zero &= ~g.mask;
            break;
        }
        default:  // H and anything else: its qubits can be 1 afterwards
            zero &= ~g.mask;
            break;
        }
    }

    uint32_t add_coef(const double* c, size_t ncomplex)
    {
        const uint32_t off = (uint32_t)(plan.coef.size() / 2);
        plan.coef.insert(plan.coef.end(), c, c + 2 * ncomplex);
        return off;
    }

    // Encodes gate g as a register op on the micro bits `bits` (ascending).
    // LDS byte offset of layout column `col`: the kernels' XOR swizzle
    // (low W bits ^= next W bits, W = 4 for 16-byte and 5 for 8-byte
    // amplitudes) times the amplitude size.
    uint32_t lds_bytes(uint32_t col) const
    {
        const uint32_t w = opt.amp_bytes == 16 ? 4u : 5u;
        return (col ^ ((col >> w) & ((1u << w) - 1u))) * (uint32_t)opt.amp_bytes;
    }

    // Order of the non-register tile bits of a register step over the thread
    // index.  A bank row is 256 bytes (16 complex128 or 32 complex64 slots);
    // lanes 0..g-1 of a wave (g = slots per row) read one row's worth of
    // addresses per LDS cycle.  Their addresses differ by XORs of the first
    // log2(g) thread-bit columns, so choosing columns whose slot indices
    // (swizzled index mod g) are linearly independent over GF(2) makes those
    // lanes hit distinct slots.  The remaining bits follow in ascending order.
    std::vector<int> lane_order(const std::vector<int>& rest, const uint16_t* cols) const
    {
        const int lg = opt.amp_bytes == 16 ? 4 : 5;
        const uint32_t mask = (1u << lg) - 1u;
        std::vector<int> order, left;
        uint32_t basis[8] = {0};  // basis[i]: vector with leading bit i
        auto reduce = [&](uint32_t v) {
            for (int i = lg - 1; i >= 0; i--)
                if (((v >> i) & 1u) && basis[i]) v ^= basis[i];
            return v;
        };
        for (int b : rest) {
            const uint32_t slot = (lds_bytes(cols[b]) / (uint32_t)opt.amp_bytes) & mask;
            const uint32_t r = (int)order.size() < lg ? reduce(slot) : 0u;
            if (r) {
                int hi = 31 - __builtin_clz(r);
                basis[hi] = r;
                order.push_back(b);
            } else {
                left.push_back(b);
            }
        }
        order.insert(order.end(), left.begin(), left.end());
        return order;
    }

    void fill_store_bytes(PassArgs& pa) const
    {
        for (int b = 0; b < 16; b++) pa.store_byte[b] = lds_bytes(pa.store_cols[b]);
    }

    // Direct first / last micro-step of a gate pass (see PassArgs): the state
    // offset of each thread-index and register bit of the step.  The first
    // step reads LDS index P, which the plain fill loads from tile index P
    // (the tile starts in load order); the last step writes LDS index P,
    // which the plain drain stores as tile index L^-1(P) for the final
    // layout L.  Only taken when the low qubits stay on lane bits 0..low-1
    // in order (coalesced runs of 2^low amplitudes, as in the plain fill and
    // drain).
    void fill_direct(PassArgs& pa) const
    {
        pa.fuse_first = pa.fuse_last = 0;
        std::fill(pa.first_g, pa.first_g + 16, 0ull);
        std::fill(pa.last_g, pa.last_g + 16, 0ull);
        if (!opt.direct || pa.micro_count == 0) return;
        const int t = plan.tile_bits, nb = opt.micro_bits, nt = t - nb;
        const int low = std::min(opt.low_qubits, nt);
        auto offset = [&](uint32_t y, const uint8_t* pos) {
            uint64_t g = 0;
            for (int j = 0; j < t; j++)
                if ((y >> j) & 1u) g ^= 1ull << pos[j];
            return g;
        };
        auto coalesced = [&](const uint64_t* g) {
            for (int i = 0; i < low; i++)
                if (g[i] != (1ull << i)) return false;
            return true;
        };
        uint64_t g[16];
        const MicroStep& f = plan.micro[pa.micro_begin];
        if (f.kind == SK_REG) {
            std::fill(g, g + 16, 0ull);
            for (int i = 0; i < nt; i++) g[i] = offset(f.tcols[i], pa.load_pos);
            for (int k = 0; k < nb; k++) g[12 + k] = offset(f.pcols[k], pa.load_pos);
            if (coalesced(g)) {
                std::copy(g, g + 16, pa.first_g);
                pa.fuse_first = 1;
            }
        }
        const MicroStep& l = plan.micro[pa.micro_begin + pa.micro_count - 1];
        if (l.kind == SK_REG) {
            // inverse of the final layout, column by column: inv[j] = tile
            // index stored at LDS index 2^j (Gauss-Jordan over GF(2))
            uint32_t a[16], inv[16];
            for (int b = 0; b < t; b++) {
                a[b] = pa.store_cols[b];
                inv[b] = 1u << b;
            }
            for (int j = 0; j < t; j++) {
                int piv = -1;
                for (int b = j; b < t; b++)
                    if ((a[b] >> j) & 1u) {
                        piv = b;
                        break;
                    }
                if (piv < 0) return;  // not invertible: keep the plain drain
                std::swap(a[j], a[piv]);
                std::swap(inv[j], inv[piv]);
                for (int b = 0; b < t; b++)
                    if (b != j && ((a[b] >> j) & 1u)) {
                        a[b] ^= a[j];
                        inv[b] ^= inv[j];
                    }
            }
            // now a[j] = 2^j and inv[j] = the tile index whose image is 2^j
            auto tile_index = [&](uint32_t p) {
                uint32_t y = 0;
                for (int j = 0; j < t; j++)
                    if ((p >> j) & 1u) y ^= inv[j];
                return y;
            };
            std::fill(g, g + 16, 0ull);
            for (int i = 0; i < nt; i++) g[i] = offset(tile_index(l.tcols[i]), pa.store_pos);
            for (int k = 0; k < nb; k++) g[12 + k] = offset(tile_index(l.pcols[k]), pa.store_pos);
            if (coalesced(g)) {
                std::copy(g, g + 16, pa.last_g);
                pa.fuse_last = 1;
            }
        }
    }

    int micro_bit(int tile_bit, const uint8_t* bits) const
    {
        for (int j = 0; j < opt.micro_bits; j++)
            if (bits[j] == tile_bit) return j;
        return -1;
    }

    std::string add_reg_op(const PlanGate& g, const int* tile_of, const uint8_t* bits)
    {
        int mi[16];
        for (int i = 0; i < g.k; i++) {
            mi[i] = micro_bit(tile_of[g.q[i]], bits);
            if (mi[i] < 0) return "internal: gate qubit outside its micro-step";
        }
        MicroOp op{};
        switch (g.id) {
        case 2: op.code = (uint16_t)(OK_H << 8 | mi[0]); break;
        case 3: op.code = (uint16_t)(OK_CNOT << 8 | (mi[0] * 4 + mi[1])); break;
        case 4: {
            const double ang = ((double)2 * 3.1415926535897932384626433) / (double)(1ull << g.optarg);
            const double c[2] = {std::cos(ang), std::sin(ang)};
            op.coef = add_coef(c, 1);
            op.code = (uint16_t)(OK_CRK << 8 | (std::min(mi[0], mi[1]) * 4 + std::max(mi[0], mi[1])));
            break;
        }
        case 5:
            op.code = (uint16_t)(OK_TOFF << 8 | (mi[2] * 16 + std::min(mi[0], mi[1]) * 4 + std::max(mi[0], mi[1])));
            break;
        case 0: {
            const double* m = prog.mat.data() + 2 * g.moff;
            if (g.k == 1 && m[2] == 0.0 && m[3] == 0.0 && m[4] == 0.0 && m[5] == 0.0) {
                const double d[4] = {m[0], m[1], m[6], m[7]};
                op.coef = add_coef(d, 2);
                op.code = (uint16_t)(OK_PHASE << 8 | mi[0]);
                break;
            }
            int order[16];
            for (int i = 0; i < g.k; i++) order[i] = i;
            std::sort(order, order + g.k, [&](int a, int b) { return mi[a] < mi[b]; });
            std::vector<double> pm = permute_matrix(m, g.k, order);
            op.coef = add_coef(pm.data(), pm.size() / 2);
            int mask = 0;
            for (int i = 0; i < g.k; i++) mask |= 1 << mi[i];
            op.code = (uint16_t)(OK_DENSE << 8 | mask);
            break;
        }
        default:
            return "unsupported gate id " + std::to_string(g.id);
        }
        plan.ops.push_back(op);
        return "";
    }

    // A dense 2^k x 2^k complex matrix (row-major, interleaved) as the lane-
    // ordered A operand of the SK_MFMA step (kernels.hip mfma_dense):
    // A[(rb*D + s)*64 + lane] = A'[rb*32 + lane%32][2s + lane/32] with
    // A'[i][2g] = Mr[i][g], A'[i][2g+1] = -Mi[i][g] for i < D (real rows) and
    // A'[D+i][2g] = Mi[i][g], A'[D+i][2g+1] = Mr[i][g] (imaginary rows).
    uint32_t add_mfma_operand(const double* m, int k)
    {
        const int D = 1 << k, RB = 2 * D / 32;
        std::vector<double> a((size_t)RB * D * 64);
        for (int rb = 0; rb < RB; rb++)
            for (int st = 0; st < D; st++)
                for (int lane = 0; lane < 64; lane++) {
                    const int i = rb * 32 + lane % 32, kk = 2 * st + lane / 32;
                    const int g = kk / 2, p = kk % 2;
                    const int row = i % D;
                    const double mr = m[2 * ((size_t)row * D + g)], mi = m[2 * ((size_t)row * D + g) + 1];
                    double v;
                    if (i < D) v = p == 0 ? mr : -mi;
                    else v = p == 0 ? mi : mr;
                    a[(size_t)(rb * D + st) * 64 + lane] = v;
                }
        const uint32_t off = (uint32_t)plan.raw.size();
        for (double v : a) {
            const float f = (float)v;
            uint32_t u;
            std::memcpy(&u, &f, 4);
            plan.raw.push_back(u);
        }
        return off;
    }

    static uint16_t bf16_rn(float f)
    {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
    }

    // The same real embedding as add_mfma_operand, split into bf16 hi + lo
    // parts and laid out for v_mfma_f32_32x32x16_bf16: for row block rb,
    // K-block t (16 contraction indices) and part (0 hi, 1 lo), lane l holds
    // A'[rb*32 + l%32][16t + 8(l/32) + j], j = 0..7, as four 32-bit words.
    uint32_t add_bf16_operand(const double* m, int k)
    {
        const int D = 1 << k, RB = 2 * D / 32, T = 2 * D / 16;
        const uint32_t off = (uint32_t)plan.raw.size();
        plan.raw.resize(plan.raw.size() + (size_t)RB * T * 2 * 64 * 4);
        for (int rb = 0; rb < RB; rb++)
            for (int t = 0; t < T; t++)
                for (int lane = 0; lane < 64; lane++) {
                    uint16_t hv[8], lv[8];
                    for (int j = 0; j < 8; j++) {
                        const int i = rb * 32 + lane % 32, kk = 16 * t + 8 * (lane / 32) + j;
                        const int g = kk / 2, p = kk % 2, row = i % D;
                        const double mr = m[2 * ((size_t)row * D + g)], mi = m[2 * ((size_t)row * D + g) + 1];
                        const double v = i < D ? (p == 0 ? mr : -mi) : (p == 0 ? mi : mr);
                        hv[j] = bf16_rn((float)v);
                        float hf;
                        const uint32_t hb = (uint32_t)hv[j] << 16;
                        std::memcpy(&hf, &hb, 4);
                        lv[j] = bf16_rn((float)(v - (double)hf));
                    }
                    for (int part = 0; part < 2; part++) {
                        const uint16_t* src = part ? lv : hv;
                        uint32_t* dst = plan.raw.data() + off + ((size_t)((rb * T + t) * 2 + part) * 64 + lane) * 4;
                        for (int w = 0; w < 4; w++) dst[w] = (uint32_t)src[2 * w] | ((uint32_t)src[2 * w + 1] << 16);
                    }
                }
        return off;
    }

    // The matrix of gate g on the window bits `wb` (tile bits, ascending),
    // left-multiplied into `u` (2^w x 2^w, row-major interleaved).
    void compose_into(std::vector<double>& u, int w, const uint8_t* wb, const PlanGate& g, const int* tile_of)
    {
        const int dim = 1 << w;
        int pos[16];
        for (int i = 0; i < g.k; i++)
            for (int j = 0; j < w; j++)
                if (wb[j] == tile_of[g.q[i]]) pos[i] = j;
        // gate matrix in its own basis
        const int gd = 1 << g.k;
        std::vector<double> gm(2 * (size_t)gd * gd, 0.0);
        auto set = [&](int r, int c, double re, double im) {
            gm[2 * ((size_t)r * gd + c)] = re;
            gm[2 * ((size_t)r * gd + c) + 1] = im;
        };
        if (g.id == 0) {
            std::copy(prog.mat.data() + 2 * g.moff, prog.mat.data() + 2 * g.moff + gm.size(), gm.begin());
        } else if (g.id == 2) {
            const double h = 0.7071067811865475;
            set(0, 0, h, 0); set(0, 1, h, 0); set(1, 0, h, 0); set(1, 1, -h, 0);
        } else if (g.id == 3) {  // (target, control): bit 0 target, bit 1 control
            set(0, 0, 1, 0); set(1, 1, 1, 0); set(2, 3, 1, 0); set(3, 2, 1, 0);
        } else if (g.id == 4) {
            const double ang = ((double)2 * 3.1415926535897932384626433) / (double)(1ull << g.optarg);
            set(0, 0, 1, 0); set(1, 1, 1, 0); set(2, 2, 1, 0); set(3, 3, std::cos(ang), std::sin(ang));
        } else {  // TOFFOLI (c0, c1, target)
            for (int x = 0; x < 8; x++) set(x, x == 3 ? 7 : (x == 7 ? 3 : x), 1, 0);
        }
        std::vector<double> out(u.size(), 0.0);
        for (int r = 0; r < dim; r++) {
            int rs = 0;
            for (int i = 0; i < g.k; i++) rs |= ((r >> pos[i]) & 1) << i;
            int rest = r;
            for (int i = 0; i < g.k; i++) rest &= ~(1 << pos[i]);
            for (int cs = 0; cs < gd; cs++) {
                const double gr = gm[2 * ((size_t)rs * gd + cs)], gi = gm[2 * ((size_t)rs * gd + cs) + 1];
                if (gr == 0.0 && gi == 0.0) continue;
                int mid = rest;
                for (int i = 0; i < g.k; i++) mid |= ((cs >> i) & 1) << pos[i];
                for (int c = 0; c < dim; c++) {
                    const double ur = u[2 * ((size_t)mid * dim + c)], ui = u[2 * ((size_t)mid * dim + c) + 1];
                    out[2 * ((size_t)r * dim + c)] += gr * ur - gi * ui;
                    out[2 * ((size_t)r * dim + c) + 1] += gr * ui + gi * ur;
                }
            }
        }
        u.swap(out);
    }

    void fill_mfma_cols(MicroStep& ms, const uint16_t* cols)
    {
        const int t = plan.tile_bits;
        uint32_t gmask = 0;
        for (int i = 0; i < ms.nbits; i++) {
            ms.gcols[i] = cols[ms.bits[i]];
            gmask |= 1u << ms.bits[i];
        }
        int j = 0;
        for (int b = 0; b < t; b++)
            if (!((gmask >> b) & 1u) && j < 12) ms.tcols[j++] = cols[b];
    }

    // A run of 1-qubit gates on one micro bit, composed in FP64.
    struct Held {
        int count = 0;
        bool single_h = false;
        double m[8] = {0};  // 2x2 complex, row-major
    };


    void hold(Held& h, const PlanGate& g)
    {
        double gm[8];
        if (g.id == 2) {
            const double s = 0.7071067811865475;
            const double hm[8] = {s, 0, s, 0, s, 0, -s, 0};
            std::copy(hm, hm + 8, gm);
        } else {
            std::copy(prog.mat.data() + 2 * g.moff, prog.mat.data() + 2 * g.moff + 8, gm);
        }
        if (h.count == 0) {
            std::copy(gm, gm + 8, h.m);
            h.single_h = g.id == 2;
        } else {
            double out[8];
            for (int r = 0; r < 2; r++)
                for (int c = 0; c < 2; c++) {
                    double re = 0, im = 0;
                    for (int k = 0; k < 2; k++) {
                        const double ar = gm[2 * (r * 2 + k)], ai = gm[2 * (r * 2 + k) + 1];
                        const double br = h.m[2 * (k * 2 + c)], bi = h.m[2 * (k * 2 + c) + 1];
                        re += ar * br - ai * bi;
                        im += ar * bi + ai * br;
                    }
                    out[2 * (r * 2 + c)] = re;
                    out[2 * (r * 2 + c) + 1] = im;
                }
            std::copy(out, out + 8, h.m);
            h.single_h = false;
        }
        h.count++;
    }

    // Euler form of a 2x2 unitary u (row-major, interleaved):
    // u = diag(p0, p1) * [[x, -iy], [-iy, x]] * diag(1, q) with x = |u00| and
    // y real.  q is set to exactly 1 (has_q false) when it differs from 1 by at
    // most 1e-15, as for RZ * RX.  Fails when a zero entry makes the phases
    // undefined or when the product differs from u by more than 1e-14 (a
    // matrix that is not unitary).
    static bool euler_form(const double* u, double& x, double& y, std::complex<double> (&p)[2],
                           std::complex<double>& q, bool& has_q)
    {
        using cd = std::complex<double>;
        const cd u00(u[0], u[1]), u01(u[2], u[3]), u10(u[4], u[5]), u11(u[6], u[7]);
        const cd I(0.0, 1.0);
        x = std::abs(u00);
        if (!(x > 1e-150) || !(std::abs(u01) > 1e-150)) return false;
        p[0] = u00 / x;
        const cd yq = I * u01 / p[0];  // u01 = -i p0 y q
        if (std::abs(yq.imag()) <= 1e-15 * std::abs(yq)) {
            y = yq.real();
            q = 1.0;
            has_q = false;
        } else {
            y = std::abs(yq);
            q = yq / y;
            has_q = true;
        }
        p[1] = I * u10 / y;  // u10 = -i p1 y
        const cd r00 = p[0] * x, r01 = -I * p[0] * y * q, r10 = -I * p[1] * y, r11 = p[1] * x * q;
        const double err = std::max(std::max(std::abs(r00 - u00), std::abs(r01 - u01)),
                                    std::max(std::abs(r10 - u10), std::abs(r11 - u11)));
        return err <= 1e-14;
    }

    // The held dense gates of the micro bits in `kron` as one Euler-form
    // Kronecker op (OK_EKRON), when every one has an Euler form and the op
    // is cheaper than the plain Kronecker op: 4 instructions per amplitude
    // per bit plus 4 per diagonal table, against 8 per bit.
    bool emit_ekron(Held* held, int kron)
    {
        using cd = std::complex<double>;
        const int K = __builtin_popcount(kron);
        if (K < 2) return false;
        double xs[4], ys[4];
        cd ps[4][2], qs[4];
        bool any_q = false;
        int j = 0;
        for (int b = 0; b < 4; b++) {
            if (!((kron >> b) & 1)) continue;
            bool hq = false;
            if (!euler_form(held[b].m, xs[j], ys[j], ps[j], qs[j], hq)) return false;
            any_q = any_q || hq;
            j++;
        }
        if (any_q && K < 3) return false;
        MicroOp op{};
        op.coef = (uint32_t)(plan.coef.size() / 2);
        for (int i = 0; i < K; i++) {
            const double c[2] = {xs[i], ys[i]};
            add_coef(c, 1);
        }
        for (int e = 0; e < (1 << K); e++) {
            cd t = 1.0;
            for (int i = 0; i < K; i++) t *= ps[i][(e >> i) & 1];
            const double c[2] = {t.real(), t.imag()};
            add_coef(c, 1);
        }
        if (any_q)
            for (int e = 0; e < (1 << K); e++) {
                cd t = 1.0;
                for (int i = 0; i < K; i++)
                    if ((e >> i) & 1) t *= qs[i];
                const double c[2] = {t.real(), t.imag()};
                add_coef(c, 1);
            }
        for (int b = 0; b < 4; b++)
            if ((kron >> b) & 1) held[b] = Held();
        op.code = (uint16_t)(OK_EKRON << 8 | kron | (any_q ? 16 : 0));
        plan.ops.push_back(op);
        return true;
    }

    // A step's lone Euler-form op on every register bit, without a right
    // diagonal, as OK_FKRON.  Each rotation is scaled by its larger entry:
    // RX_b = s_b * [[1, -i t_b], [-i t_b, 1]] with s_b = x_b, t_b = y_b / x_b
    // when |x_b| >= |y_b|, and RX_b = s_b * X_b * [[1, -i t_b], [-i t_b, 1]]
    // with s_b = -i y_b, t_b = -x_b / y_b otherwise (X_b exchanges the two
    // outputs of bit b).  With f the mask of exchanged bits,
    // D1 * prod RX_b = X_f * T * prod [[1, -i t_b], [-i t_b, 1]] where
    // T[e] = (prod s_b) * D1[e ^ f]: the kernel applies the scaled rotations
    // and T, and the step's store writes register s where s ^ f belongs.
    void factor_ekron(MicroOp& op, int nb)
    {
        using cd = std::complex<double>;
        const int full = (1 << nb) - 1;
        if ((op.code >> 8) != OK_EKRON || (op.code & 0xFF) != full) return;
        double* c = plan.coef.data() + 2 * (size_t)op.coef;
        cd scale = 1.0;
        int flip = 0;
        for (int i = 0; i < nb; i++) {
            const double x = c[2 * i], y = c[2 * i + 1];
            if (std::abs(x) >= std::abs(y)) {
                c[2 * i] = y / x;
                scale *= x;
            } else {
                c[2 * i] = -x / y;
                scale *= cd(0.0, -y);
                flip |= 1 << i;
            }
            c[2 * i + 1] = 0.0;
        }
        double* t = c + 2 * nb;
        std::vector<cd> d1((size_t)1 << nb);
        for (int e = 0; e <= full; e++) d1[(size_t)e] = cd(t[2 * e], t[2 * e + 1]);
        for (int e = 0; e <= full; e++) {
            const cd v = scale * d1[(size_t)(e ^ flip)];
            t[2 * e] = v.real();
            t[2 * e + 1] = v.imag();
        }
        op.code = (uint16_t)(OK_FKRON << 8 | full | flip << 4);
    }

    // Emits the held gates of the micro bits in `mask`: a lone Hadamard keeps
    // Kate's Hadamard arithmetic, a diagonal becomes a phase op, the rest go
    // out together as one Kronecker op (in Euler form where that is cheaper).
    void flush(Held* held, int mask)
    {
        int kron = 0;
        for (int b = 0; b < 4; b++) {
            if (!((mask >> b) & 1) || held[b].count == 0) continue;
            Held& h = held[b];
            MicroOp op{};
            if (h.count == 1 && h.single_h) {
                op.code = (uint16_t)(OK_H << 8 | b);
                plan.ops.push_back(op);
                h = Held();
            } else if (h.m[2] == 0.0 && h.m[3] == 0.0 && h.m[4] == 0.0 && h.m[5] == 0.0) {
                const double d[4] = {h.m[0], h.m[1], h.m[6], h.m[7]};
                op.coef = add_coef(d, 2);
                op.code = (uint16_t)(OK_PHASE << 8 | b);
                plan.ops.push_back(op);
                h = Held();
            } else {
                kron |= 1 << b;
            }
        }
        if (!kron) return;
        if (opt.euler && emit_ekron(held, kron)) return;
        MicroOp op{};
        op.coef = (uint32_t)(plan.coef.size() / 2);
        for (int b = 0; b < 4; b++)
            if ((kron >> b) & 1) {
                add_coef(held[b].m, 4);
                held[b] = Held();
            }
        op.code = (uint16_t)((__builtin_popcount(kron) == 1 ? OK_DENSE : OK_KRON) << 8 | kron);
        plan.ops.push_back(op);
    }

    // Packs the gates of one pass (`gl`, program order) into micro-steps.
    std::string pack_micro(const std::vector<int>& gl, const int* tile_of, PassArgs& pa)
    {
        const int t = plan.tile_bits;
        const int nb = opt.micro_bits;
        std::vector<int> rem = gl;
        pa.micro_begin = (uint32_t)plan.micro.size();
        uint16_t cols[16];
        for (int b = 0; b < 16; b++) cols[b] = (uint16_t)(1u << b);
        // A CNOT on tile bits (target tb, control cb) maps tile index y to
        // y ^ (y_cb << tb); the amplitude of index y then sits where y's image
        // sat, so the layout column of the control absorbs the target's.
        auto layout_cnot = [&](const PlanGate& g) {
            const int tb = tile_of[g.q[0]], cb = tile_of[g.q[1]];
            cols[cb] = (uint16_t)(cols[cb] ^ cols[tb]);
        };
        while (!rem.empty()) {
            std::vector<uint64_t> tm(rem.size());
            std::vector<uint8_t> ok(rem.size()), lazy(rem.size());
            for (size_t i = 0; i < rem.size(); i++) {
                const PlanGate& g = prog.gates[(size_t)rem[i]];
                uint64_t m = 0;
                for (int j = 0; j < g.k; j++) m |= 1ull << tile_of[g.q[j]];
                tm[i] = m;
                ok[i] = g.k <= nb && !(g.id == 0 && opt.mfma && g.k == 4);
                lazy[i] = g.id == 3;
            }
            if (opt.mfma_fuse) {
                // compose every gate of a window of up to 5 tile bits into
                // one dense gate for the matrix cores
                std::vector<uint8_t> ok5(rem.size()), none(rem.size(), 0);
                for (size_t i = 0; i < rem.size(); i++) ok5[i] = prog.gates[(size_t)rem[i]].k <= 5;
                uint64_t set = choose_set(tm, ok5, 5, 0, ~0ull);
                std::vector<int> acc;
                sweep(tm, ok5, set, ~0ull, nullptr, &acc);
                if (acc.empty()) return "internal: empty window";
                int w = popc(set);
                if (w < 4) {
                    set = pad_set(set, 4, t);
                    w = 4;
                }
                MicroStep ms{};
                std::copy(cols, cols + 16, ms.cols);
                ms.kind = opt.mfma_bf16 ? SK_MFMA_BF16 : SK_MFMA;
                ms.nbits = (uint8_t)w;
                int nb = 0;
                for (int b = 0; b < t; b++)
                    if ((set >> b) & 1ull) ms.bits[nb++] = (uint8_t)b;
                std::vector<double> u(2 * (size_t)(1 << w) * (1 << w), 0.0);
                for (int d = 0; d < (1 << w); d++) u[2 * ((size_t)d * (1 << w) + d)] = 1.0;
                for (int a : acc) compose_into(u, w, ms.bits, prog.gates[(size_t)rem[(size_t)a]], tile_of);
                ms.coef = opt.mfma_bf16 ? add_bf16_operand(u.data(), w) : add_mfma_operand(u.data(), w);
                fill_mfma_cols(ms, cols);
                plan.micro.push_back(ms);
                std::vector<int> next;
                size_t ai = 0;
                for (size_t i = 0; i < rem.size(); i++) {
                    if (ai < acc.size() && acc[ai] == (int)i) {
                        ai++;
                        continue;
                    }
                    next.push_back(rem[i]);
                }
                rem.swap(next);
                continue;
            }
            const uint64_t set = choose_micro_set(tm, ok, lazy, nb);
            std::vector<int> acc, late;
            sweep_micro(tm, ok, lazy, set, nullptr, &acc, &late);
            MicroStep ms{};
            std::copy(cols, cols + 16, ms.cols);
            if (acc.empty() && !late.empty()) {
                // only layout CNOTs are ready: apply them, no data moves
                for (int a : late) layout_cnot(prog.gates[(size_t)rem[(size_t)a]]);
                std::vector<int> next;
                size_t li = 0;
                for (size_t i = 0; i < rem.size(); i++) {
                    if (li < late.size() && late[li] == (int)i) {
                        li++;
                        continue;
                    }
                    next.push_back(rem[i]);
                }
                rem.swap(next);
                continue;
            }
            if (acc.empty()) {
                // the first remaining gate is one the register path does not take
                const PlanGate& g = prog.gates[(size_t)rem[0]];
                if (g.id != 0) return "internal: non-dense gate left out of micro-steps";
                const bool matrix = (opt.mfma && g.k >= 4 && g.k <= 5) || (opt.mfma5 && g.k == 5);
                ms.kind = matrix ? (opt.mfma_bf16 ? SK_MFMA_BF16 : SK_MFMA) : SK_GENERIC;
                ms.nbits = (uint8_t)g.k;
                for (int i = 0; i < g.k; i++) ms.bits[i] = (uint8_t)tile_of[g.q[i]];
                if (ms.kind == SK_MFMA || ms.kind == SK_MFMA_BF16) {
                    ms.coef = ms.kind == SK_MFMA ? add_mfma_operand(prog.mat.data() + 2 * g.moff, g.k)
                                                 : add_bf16_operand(prog.mat.data() + 2 * g.moff, g.k);
                    fill_mfma_cols(ms, cols);
                } else {
                    ms.coef = add_coef(prog.mat.data() + 2 * g.moff, (size_t)1 << (2 * g.k));
                    if (g.k == 5) {
                        // the register-blocked 5-qubit step reads the matrix
                        // by columns: its transpose follows the matrix
                        const double* u = prog.mat.data() + 2 * g.moff;
                        std::vector<double> t(2 * 1024);
                        for (int r = 0; r < 32; r++)
                            for (int c = 0; c < 32; c++) {
                                t[2 * (size_t)(c * 32 + r)] = u[2 * (size_t)(r * 32 + c)];
                                t[2 * (size_t)(c * 32 + r) + 1] = u[2 * (size_t)(r * 32 + c) + 1];
                            }
                        add_coef(t.data(), 1024);
                    }
                }
                plan.micro.push_back(ms);
                rem.erase(rem.begin());
                continue;
            }
            // pad with tile bits above the low qubits first: the low qubits
            // then stay on lane bits 0..low-1, which a direct first / last
            // step needs for coalesced HBM runs
            uint64_t full = set;
            for (int b = std::min(opt.low_qubits, t); b < t && popc(full) < nb; b++) full |= 1ull << b;
            full = pad_set(full, nb, t);
            ms.kind = SK_REG;
            ms.nbits = (uint8_t)nb;
            int nb = 0;
            for (int b = 0; b < t; b++)
                if ((full >> b) & 1ull) ms.bits[nb++] = (uint8_t)b;
            ms.op_begin = (uint32_t)plan.ops.size();
            {
                // 1-qubit gates are held per micro bit and multiplied together
                // (FP64, later gate on the left) until a gate on more qubits
                // touches the bit; held gates of several bits go out as one op.
                Held held[4];
                for (int a : acc) {
                    const PlanGate& g = prog.gates[(size_t)rem[(size_t)a]];
                    int bits[16];
                    for (int i = 0; i < g.k; i++) bits[i] = micro_bit(tile_of[g.q[i]], ms.bits);
                    if (g.k == 1) {
                        hold(held[bits[0]], g);
                        continue;
                    }
                    int touched = 0;
                    for (int i = 0; i < g.k; i++) touched |= 1 << bits[i];
                    flush(held, touched);
                    std::string e = add_reg_op(g, tile_of, ms.bits);
                    if (!e.empty()) return e;
                }
                flush(held, (1 << nb) - 1);
            }
            ms.op_count = (uint32_t)(plan.ops.size() - ms.op_begin);
            if (ms.op_count) {
                ms.op0_code = plan.ops[ms.op_begin].code;
                ms.coef = plan.ops[ms.op_begin].coef;
            }
            {
                std::vector<int> rest;
                for (int b = 0; b < t; b++) {
                    bool reg = false;
                    for (int k = 0; k < nb; k++) reg = reg || b == ms.bits[k];
                    if (!reg) rest.push_back(b);
                }
                const std::vector<int> order = lane_order(rest, ms.cols);
                int ti = 0;
                for (int b : order)
                    if (ti < 12) ms.tcols[ti++] = ms.cols[b];
                for (int k = 0; k < nb; k++) ms.pcols[k] = ms.cols[ms.bits[k]];
                for (int i = 0; i < 12; i++) ms.tbyte[i] = lds_bytes(ms.tcols[i]);
                for (int k = 0; k < 4; k++) ms.pbyte[k] = lds_bytes(ms.pcols[k]);
            }
            plan.micro.push_back(ms);
            for (int a : late) layout_cnot(prog.gates[(size_t)rem[(size_t)a]]);
            std::vector<int> done(acc);
            done.insert(done.end(), late.begin(), late.end());
            std::sort(done.begin(), done.end());
            std::vector<int> next;
            size_t ai = 0;
            for (size_t i = 0; i < rem.size(); i++) {
                if (ai < done.size() && done[ai] == (int)i) {
                    ai++;
                    continue;
                }
                next.push_back(rem[i]);
            }
            rem.swap(next);
        }
        std::copy(cols, cols + 16, pa.store_cols);
        fill_store_bytes(pa);
        pa.micro_count = (uint32_t)(plan.micro.size() - pa.micro_begin);
        pa.has_wide = 0;
        pa.light = 1;
        for (uint32_t m = pa.micro_begin; m < plan.micro.size(); m++) {
            const MicroStep& ms = plan.micro[m];
            if (ms.kind != SK_REG) pa.has_wide = 1;
            for (uint32_t o = ms.op_begin; ms.kind == SK_REG && o < ms.op_begin + ms.op_count; o++)
                if (!light_op(plan.ops[o].code)) pa.light = 0;
        }
        if (pa.has_wide) pa.light = 0;
        // the light kernels take a step's lone Euler-form op on every
        // register bit in the scaled form
        for (uint32_t m = pa.micro_begin; opt.factor && pa.light && m < plan.micro.size(); m++) {
            MicroStep& ms = plan.micro[m];
            if (ms.op_count != 1) continue;
            factor_ekron(plan.ops[ms.op_begin], ms.nbits);
            ms.op0_code = plan.ops[ms.op_begin].code;
        }
        fill_direct(pa);
        return "";
    }

    // Ops of the light kernel variant (PassArgs::light): Euler-form Kronecker
    // ops without a right diagonal, dense 1-qubit ops, phases, Hadamards,
    // CNOTs.
    static bool light_op(uint16_t code)
    {
        const int kind = code >> 8, var = code & 0xFF;
        switch (kind) {
        case OK_EKRON: return (var & 16) == 0;
        case OK_FKRON: return true;
        case OK_DENSE: return __builtin_popcount(var) == 1;
        case OK_PHASE:
        case OK_H:
        case OK_CNOT: return true;
        default: return false;
        }
    }

    // Fills the position lists of a pass whose tile holds the qubits of `set`.
    void positions(uint64_t set, PassArgs& pa, int* tile_of)
    {
        const int n = plan.nq_int;
        int tb = 0, rb = 0;
        for (int q = 0; q < n; q++) {
            if ((set >> q) & 1ull) {
                tile_of[q] = tb;
                pa.load_pos[tb] = pa.store_pos[tb] = (uint8_t)q;
                tb++;
            } else {
                tile_of[q] = -1;
                pa.rest_pos[rb++] = (uint8_t)q;
            }
        }
        pa.tile_bits = (uint8_t)tb;
        pa.nrest = (uint8_t)rb;
    }

    std::string gate_passes(const std::vector<int>& seg)
    {
        const int t = plan.tile_bits, n = plan.nq_int;
        const int low = std::min(opt.low_qubits, t);
        const uint64_t low_mask = low >= 64 ? ~0ull : ((1ull << low) - 1ull);
        std::vector<int> rem = seg;
        while (!rem.empty()) {
            std::vector<uint64_t> masks(rem.size());
            std::vector<uint8_t> ok(rem.size(), 1);
            uint64_t universe = 0;
            for (size_t i = 0; i < rem.size(); i++) {
                masks[i] = prog.gates[(size_t)rem[i]].mask;
                universe |= masks[i];
            }
            uint64_t start = low_mask;
            if (popc(low_mask | masks[0]) > t) start = pad_set(masks[0], std::min(t, popc(masks[0]) + low), n) & ~0ull;
            if (popc(start) > t) return "gate larger than the tile";
            const uint64_t set = pad_set(choose_set(masks, ok, t, start, universe), t, n);
            std::vector<int> acc;
            sweep(masks, ok, set, universe, nullptr, &acc);
            if (acc.empty()) return "internal: empty pass";
            PassArgs pa{};
            pa.micro_bits = (uint8_t)opt.micro_bits;
            int tile_of[kMaxQubits];
            positions(set, pa, tile_of);
            std::vector<int> gl;
            for (int a : acc) gl.push_back(rem[(size_t)a]);
            std::string e = pack_micro(gl, tile_of, pa);
            if (!e.empty()) return e;
            plan.pass_zero.push_back(remapped ? 0ull : zero);
            if (!remapped) plan.support_passes = (int)plan.passes.size() + 1;
            for (int gi : gl) track(prog.gates[(size_t)gi]);
            plan.passes.push_back(pa);
            plan.pass_is_remap.push_back(0);
            plan.gate_passes++;
            std::vector<int> next;
            size_t ai = 0;
            for (size_t i = 0; i < rem.size(); i++) {
                if (ai < acc.size() && acc[ai] == (int)i) {
                    ai++;
                    continue;
                }
                next.push_back(rem[i]);
            }
            rem.swap(next);
        }
        return "";
    }

    // A swap instruction (Kate's swapCommand, HIPcode/simulator.hpp:751-759):
    // its transpositions of qubit positions, applied in order, become the
    // store positions of remap passes.
    std::string remap_passes(const std::vector<int>& pairs)
    {
        const int t = plan.tile_bits, n = plan.nq_int;
        const int low = std::min(opt.low_qubits, t);
        size_t p = 0;
        while (p < pairs.size()) {
            uint64_t set = (1ull << low) - 1ull;
            size_t end = p;
            while (end + 1 < pairs.size()) {
                uint64_t s2 = set | (1ull << pairs[end]) | (1ull << pairs[end + 1]);
                if (popc(s2) > t) break;
                set = s2;
                end += 2;
            }
            if (end == p) return "swap pair does not fit a tile";
            set = pad_set(set, t, n);
            PassArgs pa{};
            pa.micro_bits = (uint8_t)opt.micro_bits;
            int tile_of[kMaxQubits];
            positions(set, pa, tile_of);
            int where[kMaxQubits];
            for (int q = 0; q < n; q++) where[q] = q;
            for (size_t i = p; i < end; i += 2) std::swap(where[pairs[i]], where[pairs[i + 1]]);
            for (int b = 0; b < pa.tile_bits; b++) pa.store_pos[b] = (uint8_t)where[pa.load_pos[b]];
            pa.micro_begin = (uint32_t)plan.micro.size();
            pa.micro_count = 0;
            for (int b = 0; b < 16; b++) pa.store_cols[b] = (uint16_t)(1u << b);
            fill_store_bytes(pa);
            // a remap pass writes every tile (masked reads of the zero
            // qubits); the support is not tracked through it
            plan.pass_zero.push_back(remapped ? 0ull : zero);
            remapped = true;
            zero = 0;
            plan.passes.push_back(pa);
            plan.pass_is_remap.push_back(1);
            plan.remap_passes++;
            p = end;
        }
        return "";
    }
};

}  // namespace

PassArgs zero_start_pass(const Plan& plan, size_t i, bool support, size_t n_run)
{
    PassArgs pa = plan.passes[i];
    if (i == 0) {
        pa.synth_basis = 1;
        pa.synth_index = 0;
    }
    if (support && i < plan.pass_zero.size()) {
        // Every tile that is skipped is zero in and out, and no later pass
        // reads it before a pass has written it: a rest qubit keeps its zero
        // status through a pass, and the last pass writes every tile.
        const uint64_t z = plan.pass_zero[i];
        if (i > 0) pa.zero_mask = z;
        if ((int)i < plan.support_passes && i + 1 < n_run) {
            int k = 0;
            for (int j = 0; j < pa.nrest; j++)
                if (!((z >> pa.rest_pos[j]) & 1ull)) pa.rest_pos[k++] = pa.rest_pos[j];
            pa.nrest = (uint8_t)k;
        }
    }
    return pa;
}

// Inverse of an invertible GF(2) matrix given by its columns (n <= 64):
// inv[j] = the vector whose image is 2^j.  False if it is singular.
static bool gf2_inverse(std::vector<uint64_t> a, std::vector<uint64_t>& inv)
{
    const size_t n = a.size();
    inv.assign(n, 0);
    for (size_t b = 0; b < n; b++) inv[b] = 1ull << b;
    for (size_t j = 0; j < n; j++) {
        size_t piv = n;
        for (size_t b = j; b < n; b++)
            if ((a[b] >> j) & 1ull) {
                piv = b;
                break;
            }
        if (piv == n) return false;
        std::swap(a[j], a[piv]);
        std::swap(inv[j], inv[piv]);
        for (size_t b = 0; b < n; b++)
            if (b != j && ((a[b] >> j) & 1ull)) {
                a[b] ^= a[j];
                inv[b] ^= inv[j];
            }
    }
    return true;
}

// The trailing pure-permutation passes as one linear map of state offsets
// (Plan::trail_begin, Plan::trail_inv).  A pass without micro-steps moves
// the amplitude at tile index u (rest offset r) to r ^ store(L^-1(u)), L its
// final layout.
static void plan_trailing(Plan& plan)
{
    const int n = plan.nq_int;
    size_t b = plan.passes.size();
    while (b > 0 && plan.passes[b - 1].micro_count == 0) b--;
    plan.trail_begin = b;
    plan.trail_inv.clear();
    if (b == plan.passes.size() || n > 64) {
        plan.trail_begin = plan.passes.size();
        return;
    }
    std::vector<uint64_t> img((size_t)n);
    for (int q = 0; q < n; q++) img[(size_t)q] = 1ull << q;
    for (size_t k = b; k < plan.passes.size(); k++) {
        const PassArgs& pa = plan.passes[k];
        const int t = pa.tile_bits;
        std::vector<uint64_t> cols((size_t)t), linv;
        for (int j = 0; j < t; j++) cols[(size_t)j] = pa.store_cols[j];
        if (!gf2_inverse(cols, linv)) {
            plan.trail_begin = plan.passes.size();
            return;
        }
        for (int q = 0; q < n; q++) {
            uint64_t v = img[(size_t)q], u = 0;
            for (int j = 0; j < t; j++)
                if ((v >> pa.load_pos[j]) & 1ull) {
                    u |= 1ull << j;
                    v &= ~(1ull << pa.load_pos[j]);
                }
            uint64_t y = 0;
            for (int j = 0; j < t; j++)
                if ((u >> j) & 1ull) y ^= linv[(size_t)j];
            for (int j = 0; j < t; j++)
                if ((y >> j) & 1ull) v ^= 1ull << pa.store_pos[j];
            img[(size_t)q] = v;
        }
    }
    if (!gf2_inverse(img, plan.trail_inv)) {
        plan.trail_begin = plan.passes.size();
        plan.trail_inv.clear();
    }
}

// The trailing permutation passes (together the map A, known as
// trail_inv = A^-1) folded into the last gate pass's store when its tile
// holds every qubit that A^-1 of a low-qubit offset involves; else, when
// there are two or more, as one out-of-place pass.  That pass's tile holds the low
// qubits and every qubit that A^-1 of a low-qubit offset involves, so that
// each output run of 2^low amplitudes has all its sources in one tile; the
// tile is read in runs as usual and written through A: the store index has
// the low offset bits first, then tile-bit images completing a basis of
// their span, and each store index bit k reads LDS through store_cols[k]
// (its combination of tile bits) and writes to out_tile[k].
static void make_scatter(Plan& plan, int amp_bytes)
{
    const size_t b = plan.trail_begin, np = plan.passes.size();
    if (np <= b || plan.trail_inv.empty() || plan.remap_passes) return;
    for (size_t k = b; k < np; k++)
        if (plan.pass_is_remap[k]) return;
    const int n = plan.nq_int, t = plan.tile_bits, low = std::min(plan.low_qubits, plan.tile_bits);
    if (n > 64) return;
    std::vector<uint64_t> A;
    if (!gf2_inverse(plan.trail_inv, A)) return;
    uint64_t set = (1ull << low) - 1ull;
    for (int j = 0; j < low; j++) set |= plan.trail_inv[(size_t)j];
    const uint32_t w = amp_bytes == 16 ? 4u : 5u;
    auto lds_bytes = [&](uint32_t col) { return (col ^ ((col >> w) & ((1u << w) - 1u))) * (uint32_t)amp_bytes; };
    // When the last gate pass's tile holds those qubits, that pass itself
    // stores through A (out of place), and no permutation pass remains.
    if (b > 0 && !plan.pass_is_remap[b - 1] && plan.passes[b - 1].micro_count > 0) {
        PassArgs& gp = plan.passes[b - 1];
        uint64_t spos = 0;
        for (int p = 0; p < gp.tile_bits; p++) spos |= 1ull << gp.store_pos[p];
        if ((set & ~spos) == 0 && gp.tile_bits == t) {
            auto store_coords = [&](uint64_t x) {
                uint64_t c = 0;
                for (int p = 0; p < t; p++)
                    if ((x >> gp.store_pos[p]) & 1ull) c |= 1ull << p;
                return c;
            };
            std::vector<uint64_t> out, coef, red;
            auto add = [&](uint64_t v, uint64_t c) {
                uint64_t x = v;
                for (uint64_t r : red)
                    if ((x >> __builtin_ctzll(r)) & 1ull) x ^= r;
                if (!x) return false;
                red.push_back(x);
                out.push_back(v);
                coef.push_back(c);
                return true;
            };
            bool ok = true;
            for (int j = 0; j < low && ok; j++) ok = add(1ull << j, store_coords(plan.trail_inv[(size_t)j]));
            for (int p = 0; p < t && ok && (int)out.size() < t; p++) add(A[gp.store_pos[p]], 1ull << p);
            if (ok && (int)out.size() == t) {
                ScatterArgs sa{};
                uint16_t cols[16];
                for (int k = 0; k < t; k++) {
                    uint32_t col = 0;
                    for (int p = 0; p < t; p++)
                        if ((coef[(size_t)k] >> p) & 1ull) col ^= gp.store_cols[p];
                    cols[k] = (uint16_t)col;
                    sa.out_tile[k] = out[(size_t)k];
                }
                for (int k = 0; k < t; k++) {
                    gp.store_cols[k] = cols[k];
                    gp.store_byte[k] = lds_bytes(cols[k]);
                }
                for (int i = 0; i < gp.nrest; i++) sa.out_rest[i] = A[gp.rest_pos[i]];
                gp.scatter = 1;
                gp.scatter_idx = (uint32_t)plan.scatter_args.size();
                gp.fuse_last = 0;  // the last micro-step writes LDS, the scatter store reads it
                plan.passes.resize(b);
                plan.pass_is_remap.resize(b);
                if (plan.pass_zero.size() > b) plan.pass_zero.resize(b);
                plan.gate_passes -= (int)(np - b);
                plan.support_passes = std::min(plan.support_passes, (int)b - 1);
                plan.scatter_pass = (long)(b - 1);
                plan.scatter_args.push_back(sa);
                plan.trail_begin = plan.passes.size();
                plan.trail_inv.clear();
                return;
            }
        }
    }
    if (np < b + 2 || popc(set) > t) return;  // one pass stays as it is
    set = pad_set(set, t, n);
    PassArgs pa{};
    int tb = 0, rb = 0;
    for (int q = 0; q < n; q++) {
        if ((set >> q) & 1ull) pa.load_pos[tb++] = (uint8_t)q;
        else pa.rest_pos[rb++] = (uint8_t)q;
    }
    pa.tile_bits = (uint8_t)tb;
    pa.nrest = (uint8_t)rb;
    pa.micro_bits = plan.passes[b].micro_bits;
    pa.micro_begin = (uint32_t)plan.micro.size();
    pa.micro_count = 0;
    pa.scatter = 1;
    pa.scatter_idx = (uint32_t)plan.scatter_args.size();
    auto tile_coords = [&](uint64_t x) {
        uint64_t c = 0;
        for (int k = 0; k < t; k++)
            if ((x >> pa.load_pos[k]) & 1ull) c |= 1ull << k;
        return c;
    };
    // basis of the span of the tile-bit images, starting with the low offsets
    std::vector<uint64_t> out, coef, red, redc;  // red: reduced rows (pivot = lowest set bit)
    auto add = [&](uint64_t v, uint64_t c) {
        uint64_t x = v, xc = c;
        for (size_t i = 0; i < red.size(); i++)
            if ((x >> __builtin_ctzll(red[i])) & 1ull) {
                x ^= red[i];
                xc ^= redc[i];
            }
        if (!x) return false;
        red.push_back(x);
        redc.push_back(xc);
        out.push_back(v);
        coef.push_back(c);
        return true;
    };
    for (int j = 0; j < low; j++)
        if (!add(1ull << j, tile_coords(plan.trail_inv[(size_t)j]))) return;
    for (int k = 0; k < t && (int)out.size() < t; k++) add(A[pa.load_pos[k]], 1ull << k);
    if ((int)out.size() != t) return;
    ScatterArgs sa{};
    for (int k = 0; k < t; k++) {
        const uint32_t col = (uint32_t)coef[(size_t)k];
        pa.store_cols[k] = (uint16_t)col;
        pa.store_byte[k] = lds_bytes(col);
        pa.store_pos[k] = pa.load_pos[k];
        sa.out_tile[k] = out[(size_t)k];
    }
    for (int i = 0; i < rb; i++) sa.out_rest[i] = A[pa.rest_pos[i]];
    // a run from |0...0> reads the input through the zero qubits of the first
    // merged pass; the scatter pass launches every tile (it writes the other
    // buffer, whose old contents must all be replaced)
    const bool has_zero = plan.pass_zero.size() > b;
    const uint64_t z = has_zero ? plan.pass_zero[b] : 0ull;
    plan.passes.resize(b);
    plan.pass_is_remap.resize(b);
    if (has_zero) plan.pass_zero.resize(b);
    plan.passes.push_back(pa);
    plan.pass_is_remap.push_back(0);
    if (has_zero) plan.pass_zero.push_back(z);
    plan.gate_passes -= (int)(np - b) - 1;
    plan.support_passes = std::min(plan.support_passes, (int)b);
    plan.scatter_pass = (long)b;
    plan.scatter_args.push_back(sa);
}

// Relabelled passes.  A pass whose tile qubits sit at state bits
// 0..tile_bits-1 reads whole 64 KiB blocks; one whose tile qubits are high
// bits reads 2^low-amplitude runs far apart, which the memory system serves
// more slowly (one MI355X, complex128, 30 qubits: an out-of-place copy of
// tiles over qubits {0,1,2,18..26} 8.3 ms, reading contiguous tiles and
// writing to those positions 6.7 ms).  From the first pass after which no
// qubit is still known zero (earlier passes skip tiles in place; a plan for
// any input state, PlanOptions::zero_start off, has none), every pass is
// out of place: it reads its tile from bits 0..tile_bits-1 and writes its
// amplitudes where the next pass's tile and rest qubits are placed (the last
// one to virtual order, or through the trailing permutation).  Only physical
// positions change; the arithmetic is that of the plain plan, bit for bit.
// 30 qubits depth 20, 12-qubit tiles with 3 low qubits: complex128
// 182 -> 163 ms, complex64 105 -> 98 ms.  Since the reads no longer depend
// on the low qubits, plans with fewer of them (fewer passes, shorter write
// runs) become the cheaper ones; see plan_cost.
static uint64_t permute_bits(const std::vector<int>& to, uint64_t x)
{
    uint64_t y = 0;
    while (x) {
        const int q = __builtin_ctzll(x);
        x &= x - 1;
        y |= 1ull << to[(size_t)q];
    }
    return y;
}

static void make_relabel(Plan& plan)
{
    const int n = plan.nq_int, t = plan.tile_bits;
    const size_t np = plan.passes.size();
    if (n > 64 || plan.remap_passes || np < 3 || plan.pass_zero.size() < np) return;
    size_t start = np;
    while (start > 0 && plan.pass_zero[start - 1] == 0) start--;
    if (start + 2 > np) return;
    for (size_t k = start; k < np; k++)
        if (plan.pass_is_remap[k] || plan.passes[k].tile_bits != t || (k + 1 < np && plan.passes[k].scatter)) return;
    // every pass must read and (unless it already scatters) write each state
    // bit exactly once, so that the relabelling maps are permutations
    auto covers = [&](const uint8_t* pos, const PassArgs& pa) {
        if (pa.nrest != n - t) return false;
        uint64_t seen = 0;
        for (int b = 0; b < t; b++) seen |= 1ull << pos[b];
        for (int i = 0; i < pa.nrest; i++) seen |= 1ull << pa.rest_pos[i];
        return seen == (n == 64 ? ~0ull : (1ull << n) - 1) && __builtin_popcountll(seen) == n;
    };
    for (size_t k = start; k < np; k++) {
        const PassArgs& pa = plan.passes[k];
        if (!covers(pa.load_pos, pa) || (!pa.scatter && !covers(pa.store_pos, pa))) return;
    }
    std::vector<int> P((size_t)n), Q((size_t)n);
    for (int q = 0; q < n; q++) P[(size_t)q] = q;
    for (size_t k = start; k < np; k++) {
        PassArgs& pa = plan.passes[k];
        if (k + 1 < np) {
            const PassArgs& nx = plan.passes[k + 1];
            for (int b = 0; b < t; b++) Q[nx.load_pos[b]] = b;
            for (int i = 0; i < nx.nrest; i++) Q[nx.rest_pos[i]] = t + i;
        } else {
            for (int q = 0; q < n; q++) Q[(size_t)q] = q;
        }
        if (!pa.scatter) {
            ScatterArgs sa{};
            for (int p = 0; p < t; p++) sa.out_tile[p] = 1ull << Q[pa.store_pos[p]];
            for (int i = 0; i < pa.nrest; i++) sa.out_rest[i] = 1ull << Q[pa.rest_pos[i]];
            for (int i = 0; i < 16; i++) pa.last_g[i] = permute_bits(Q, pa.last_g[i]);
            pa.scatter = 1;
            pa.scatter_idx = (uint32_t)plan.scatter_args.size();
            plan.scatter_args.push_back(sa);
        }
        for (int b = 0; b < t; b++) pa.load_pos[b] = (uint8_t)P[pa.load_pos[b]];
        for (int i = 0; i < pa.nrest; i++) pa.rest_pos[i] = (uint8_t)P[pa.rest_pos[i]];
        for (int i = 0; i < 16; i++) pa.first_g[i] = permute_bits(P, pa.first_g[i]);
        P = Q;
    }
    plan.support_passes = std::min(plan.support_passes, (int)start);
    plan.relabel_begin = (long)start;
    // execute() cannot measure through trailing passes of a relabelled plan
    plan.trail_begin = np;
    plan.trail_inv.clear();
}

// Relative HBM time of a pass whose tiles are read and written in runs of
// `run_bytes` contiguous bytes (one MI355X, 30-qubit passes: 64 B runs cost
// 1.37x, 256 B runs 0.92x the time of 128 B runs).
static double run_cost(int run_bytes)
{
    if (run_bytes <= 32) return 2.5;
    if (run_bytes <= 64) return 1.37;
    if (run_bytes <= 128) return 1.0;
    if (run_bytes <= 256) return 0.92;
    return 0.9;
}

// The same swap-free program with its gates in a canonical topological
// order: by ASAP level (one more than the deepest earlier gate on a shared
// qubit), then lowest qubit, then program order; one compute group.  Every
// order of the same dependency graph (Kate's gateScheduling only reorders
// within it, Circuit.hpp:12-98) maps to the same canonical order, so the
// light-cone packer, whose greedy choices break ties by program order, no
// longer depends on which topological order the caller passed.
static Program canonical_order(const Program& in)
{
    Program out = in;
    std::vector<int> level(64, 0);
    std::vector<std::pair<std::pair<int, int>, int>> key;
    key.reserve(in.gates.size());
    for (size_t i = 0; i < in.gates.size(); i++) {
        const PlanGate& g = in.gates[i];
        int l = 0, lo = 64;
        for (int j = 0; j < g.k; j++) {
            l = std::max(l, level[(size_t)g.q[j]]);
            lo = std::min(lo, g.q[j]);
        }
        for (int j = 0; j < g.k; j++) level[(size_t)g.q[j]] = l + 1;
        key.push_back({{l + 1, lo}, (int)i});
    }
    std::sort(key.begin(), key.end());
    out.gates.clear();
    uint64_t mask = 0;
    for (const auto& k : key) {
        out.gates.push_back(in.gates[(size_t)k.second]);
        mask |= in.gates[(size_t)k.second].mask;
    }
    out.group_end.assign(1, (int)out.gates.size());
    out.group_mask.assign(1, mask);
    out.instr_kind.assign(1, 1);
    out.instr_pairs.assign(1, std::vector<int>());
    return out;
}

// Relative HBM time of a relabelled pass, which reads whole tile blocks and
// writes runs of `run_bytes`, in the units of run_cost.  Fitted on one
// MI355X, 30-qubit layered circuit: relabelled passes of s register steps
// take about 0.5 + 0.1 s (complex128) and 0.67 + 0.11 s (complex64) times
// an in-place 128 B-run copy, for 64 B write runs as for longer ones; 32 B
// runs double some passes.
static double relabel_run_cost(int run_bytes)
{
    return run_bytes <= 32 ? 2.5 : 0.6;
}

// Relative cost of a plan: per pass the HBM cost of its run length (in
// place: reads and writes; relabelled: the writes only) times the fraction
// of its tiles that a run from |0...0> launches (a tile whose rest qubits
// include one still known zero is skipped), plus a tenth of a pass per
// micro-step.
static double plan_cost(const Plan& p, int amp_bytes)
{
    const int run = (1 << std::min(p.low_qubits, p.nq_int)) * amp_bytes;
    double c = 0.1 * (double)p.micro.size();
    for (size_t i = 0; i < p.passes.size(); i++) {
        const PassArgs& pa = p.passes[i];
        double f = 1.0;
        if ((int)i < p.support_passes && i < p.pass_zero.size())
            for (int j = 0; j < pa.nrest; j++)
                if ((p.pass_zero[i] >> pa.rest_pos[j]) & 1ull) f *= 0.5;
        // (a pass of the full op set is bound by its register steps; its
        // reads cost it what they cost in place)
        const bool relabelled = p.relabel_begin >= 0 && (long)i >= p.relabel_begin && pa.light;
        c += f * (relabelled ? relabel_run_cost(run) : run_cost(run));
    }
    return c;
}

std::string build_plan(const Program& prog, const PlanOptions& opt, Plan& plan)
{
    if (opt.split) {
        PlanOptions o = opt;
        o.split = false;
        return build_plan(split_kron_gates(prog), o, plan);
    }
    if (opt.auto_tile) {
        // complex64 by default: 13-qubit tiles (64 KiB; 8 free tile qubits
        // beside 5 low ones) when every gate pass runs on the light kernel
        // variant and the plan is no dearer; 12-qubit tiles otherwise.
        // Measured on one MI355X, layered circuits, 13 against 12 qubits:
        // 30 qubits depth 20 112 / 123 ms, 28 qubits depth 20 25.5 / 27.3 ms,
        // 26 qubits depth 30 9.4 / 9.8 ms, 30 qubits depth 40 265 / 257 ms.
        // The passes of the full op set (dense 2-qubit, CRk, Toffoli ops)
        // are compute-bound, and there four 32 KiB workgroups per CU beat
        // two of 64 KiB (32-qubit Haar brickwork: 36 passes of 13 qubits
        // 1021 ms, 44 of 12 qubits 988 ms).
        PlanOptions o = opt;
        o.auto_tile = false;
        o.tile_bits = 12;
        o.min_qubits = 12;
        std::string e = build_plan(prog, o, plan);
        if (!e.empty() || prog.nqubits < 13) return e;
        o.tile_bits = 13;
        o.min_qubits = 13;
        Plan alt;
        if (!build_plan(prog, o, alt).empty()) return "";
        bool all_light = true;
        for (size_t i = 0; i < alt.passes.size(); i++)
            if (!alt.pass_is_remap[i] && alt.passes[i].micro_count && !alt.passes[i].light) all_light = false;
        if (all_light && plan_cost(alt, opt.amp_bytes) <= plan_cost(plan, opt.amp_bytes)) plan = std::move(alt);
        return "";
    }
    if (opt.reorder) {
        PlanOptions o = opt;
        o.reorder = false;
        std::string e = build_plan(prog, o, plan);
        bool swap_free = !prog.gates.empty();
        for (int k : prog.instr_kind) swap_free = swap_free && k == 1;
        if (!e.empty() || !swap_free) return e;
        Plan alt;
        if (build_plan(canonical_order(prog), o, alt).empty() &&
            plan_cost(alt, opt.amp_bytes) < plan_cost(plan, opt.amp_bytes))
            plan = std::move(alt);
        return "";
    }
    if (opt.auto_low) {
        // More low qubits per tile give longer HBM runs but fewer free tile
        // qubits, hence more passes; the light-cone packer's pass count is
        // not monotonic in it (28-qubit layered circuit: 32 passes with 3
        // low qubits, 28 with 4).  Plan each candidate and keep the cheapest
        // by pass count x run cost + 0.1 pass per micro-step (measured: a
        // register step costs about a tenth of a 128 B-run pass).  With
        // relabelled passes the reads are whole blocks whatever the low
        // qubits, so plans with fewer low qubits and fewer passes gain
        // (30-qubit layered circuit, complex64 13-qubit tiles: 18 passes with
        // 3 low qubits 85.6 ms, 25 with 5 96.6 ms).  Relabelled complex128
        // plans keep at least 3 low qubits (128 B runs): with 2 the same
        // circuit takes 19 passes and 158 ms against 24 passes and 162 ms,
        // but each pass moves its 34.4 GB in 10.2 ms instead of 7.9 ms; the
        // longer runs are kept for the passes' nearness to their HBM bound.
        PlanOptions o = opt;
        o.auto_low = false;
        std::string first_err;
        double best = 0;
        bool have = false;
        const int lo = opt.amp_bytes == 16 && !(opt.scatter && opt.relabel) ? 2 : 3;
        for (int low = lo; low <= lo + 3; low++) {
            if (low > std::min(opt.tile_bits, kMaxQubits) - opt.micro_bits) break;
            o.low_qubits = low;
            Plan cand;
            std::string e = build_plan(prog, o, cand);
            if (!e.empty()) {
                if (first_err.empty()) first_err = e;
                continue;
            }
            const double cost = plan_cost(cand, opt.amp_bytes);
            if (!have || cost < best) {
                best = cost;
                plan = std::move(cand);
                have = true;
            }
        }
        if (!have) return first_err.empty() ? "no plan" : first_err;
        return "";
    }
    plan = Plan();
    plan.low_qubits = opt.low_qubits;
    plan.nq_int = std::max(prog.nqubits, opt.min_qubits);
    if (plan.nq_int > kMaxQubits) return "too many qubits";
    plan.tile_bits = std::min(opt.tile_bits, plan.nq_int);
    if (opt.micro_bits != 3 && opt.micro_bits != 4) return "micro-steps have 3 or 4 bits";
    if (opt.amp_bytes != 16 && opt.amp_bytes != 8) return "amplitudes have 8 or 16 bytes";
    if (opt.micro_bits != 4 && (opt.mfma || opt.mfma_fuse || opt.mfma5)) return "matrix-core steps need 4-bit micro-steps";
    Builder b{prog, opt, plan};
    b.zero = !opt.zero_start ? 0ull : plan.nq_int >= 64 ? ~0ull : ((1ull << plan.nq_int) - 1ull);
    std::vector<int> seg;
    size_t grp = 0;
    int first = 0;
    for (size_t i = 0; i < prog.instr_kind.size(); i++) {
        if (prog.instr_kind[i] == 0) {
            if (!seg.empty()) {
                std::string e = b.gate_passes(seg);
                if (!e.empty()) return e;
                seg.clear();
            }
            std::string e = b.remap_passes(prog.instr_pairs[i]);
            if (!e.empty()) return e;
        } else {
            if (grp >= prog.group_end.size()) return "more compute instructions than groups";
            for (int g = first; g < prog.group_end[grp]; g++) seg.push_back(g);
            first = prog.group_end[grp];
            grp++;
        }
    }
    if (!seg.empty()) {
        std::string e = b.gate_passes(seg);
        if (!e.empty()) return e;
    }
    plan_trailing(plan);
    if (opt.scatter) make_scatter(plan, opt.amp_bytes);
    if (opt.scatter && opt.relabel) make_relabel(plan);
    return "";
}

std::string Plan::describe(bool verbose) const
{
    std::ostringstream os;
    os << "nq_int=" << nq_int << " tile_bits=" << tile_bits << " low=" << low_qubits << " passes=" << passes.size()
       << " (gate " << gate_passes
       << ", remap " << remap_passes << ") micro=" << micro.size() << " ops=" << ops.size() << "\n";
    for (size_t i = 0; i < passes.size(); i++) {
        const PassArgs& p = passes[i];
        os << (pass_is_remap[i] ? "remap" : "pass ") << " " << i << " tile=[";
        for (int b = 0; b < p.tile_bits; b++) os << (b ? "," : "") << (int)p.load_pos[b];
        os << "]";
        if (pass_is_remap[i]) {
            os << " -> [";
            for (int b = 0; b < p.tile_bits; b++) os << (b ? "," : "") << (int)p.store_pos[b];
            os << "]";
        }
        size_t nops = 0;
        for (uint32_t m = 0; m < p.micro_count; m++) nops += micro[p.micro_begin + m].op_count;
        os << " micro=" << p.micro_count << " ops=" << nops;
        if (p.fuse_first || p.fuse_last)
            os << " direct=" << (p.fuse_first ? "first" : "") << (p.fuse_first && p.fuse_last ? "," : "")
               << (p.fuse_last ? "last" : "");
        if (p.light && p.micro_count) os << " light";
        if (p.scatter) os << ((long)i == scatter_pass ? " scatter" : " relabel");
        // state bits known zero at this pass's input in a run from |0...0>
        // (its loads skip those amplitudes)
        if (i < pass_zero.size() && pass_zero[i]) os << " zero_in=" << __builtin_popcountll(pass_zero[i]);
        os << "\n";
        if (!verbose) continue;
        static const char* kinds[] = {"reg", "generic", "mfma", "mfma_bf16"};
        static const char* okinds[] = {"dense", "phase", "h", "cnot", "crk", "toff", "kron", "ekron", "fkron"};
        for (uint32_t m = 0; m < p.micro_count; m++) {
            const MicroStep& ms = micro[p.micro_begin + m];
            os << "    step " << m << " " << kinds[ms.kind] << " bits=[";
            const int nb = ms.nbits;
            for (int b = 0; b < nb; b++) os << (b ? "," : "") << (int)ms.bits[b];
            os << "]";
            for (uint32_t o = 0; o < ms.op_count; o++) {
                const MicroOp& op = ops[ms.op_begin + o];
                os << " " << okinds[op.code >> 8] << ":" << (op.code & 0xFF);
            }
            if (ms.kind == SK_REG) {
                os << " tcols=";
                for (int i = 0; i + ms.nbits < p.tile_bits; i++) os << (i ? "," : "") << ms.tcols[i];
                os << " pcols=";
                for (int k = 0; k < 4; k++) os << (k ? "," : "") << ms.pcols[k];
            }
            os << "\n";
        }
        os << "    store_cols=";
        for (int b = 0; b < p.tile_bits; b++) os << (b ? "," : "") << p.store_cols[b];
        if (p.fuse_first) {
            os << " first_g=";
            for (int i = 0; i < 16; i++) os << (i ? "," : "") << p.first_g[i];
        }
        if (p.fuse_last) {
            os << " last_g=";
            for (int i = 0; i < 16; i++) os << (i ? "," : "") << p.last_g[i];
        }
        if (p.scatter && p.scatter_idx < scatter_args.size()) {
            const ScatterArgs& sa = scatter_args[p.scatter_idx];
            os << " out_tile=";
            for (int b = 0; b < p.tile_bits; b++) os << (b ? "," : "") << sa.out_tile[b];
            os << " out_rest=";
            for (int i = 0; i < p.nrest; i++) os << (i ? "," : "") << sa.out_rest[i];
        }
        os << "\n";
    }
    return os.str();
}

}  // namespace kate
