/*
 * kate_oracle.c -- TEST INFRASTRUCTURE ONLY (see kate_oracle.h).
 *
 * Plain-C restatement of Kate's CPU simulator.  Built with -ffp-contract=off
 * so that every complex product is rounded exactly as Kate's Complex class
 * rounds it (backendagnosticcode/basic_host_types.hpp:43-45), which makes the
 * oracle bit-identical to Kate's CPU simulator on the same program.
 */
#include "kate_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* Constants as Kate defines them (backendagnosticcode/agnosticpreprocessor.hpp:25-27). */
#define ORC_PI 3.1415926535897932384626433
#define ORC_SQRT2 1.4142135623730951
#define ORC_SQRT2INV 0.7071067811865475

/* Zero bits are inserted into `line` at each position listed (ascending) in
 * `pos`, i.e. the masks of GateComputing.hpp:91-102: the result is the base
 * index of a gate's sub-vector. */
static inline uint64_t insert_zeros(uint64_t line, const int* pos, int k)
{
    for (int i = 0; i < k; i++) {
        uint64_t low = line & ((1ull << pos[i]) - 1ull);
        line = low | ((line - low) << 1);
    }
    return line;
}

static void sort_ints(int* v, int k)
{
    for (int i = 1; i < k; i++) {
        int x = v[i], j = i - 1;
        while (j >= 0 && v[j] > x) { v[j + 1] = v[j]; j--; }
        v[j + 1] = x;
    }
}

/* computeGate, CPUcode/GateComputing.hpp:6-128. */
int orc_apply_gate(const orc_gates* gs, int g, int nlocal, double* st, const int32_t* b2l)
{
    const int id = gs->id[g];
    const int k = gs->nq[g];
    const int32_t* q = gs->qubits + gs->qoff[g];
    int lq[16], sorted[16];
    if (k < 1 || k > 16 || k > nlocal) return -1;
    for (int i = 0; i < k; i++) { lq[i] = b2l[q[i]]; sorted[i] = lq[i]; }
    sort_ints(sorted, k);
    const uint64_t cover = 1ull << (nlocal - k);

    switch (id) {
    case ORC_TOFFOLI: { /* controls q[0], q[1]; target q[2]  (:8-28) */
        if (k != 3) return -1;
        for (uint64_t line = 0; line < cover; line++) {
            uint64_t i110 = insert_zeros(line, sorted, 3) | (1ull << lq[0]) | (1ull << lq[1]);
            uint64_t i111 = i110 | (1ull << lq[2]);
            double r = st[2 * i110], m = st[2 * i110 + 1];
            st[2 * i110] = st[2 * i111]; st[2 * i110 + 1] = st[2 * i111 + 1];
            st[2 * i111] = r; st[2 * i111 + 1] = m;
        }
        return 0;
    }
    case ORC_CRK: { /* phase e^{2 pi i / 2^k} on |11>  (:29-45) */
        if (k != 2) return -1;
        for (uint64_t line = 0; line < cover; line++) {
            uint64_t i11 = insert_zeros(line, sorted, 2) | (1ull << lq[1]) | (1ull << lq[0]);
            double ang = ((double)2 * ORC_PI) / (double)(1ull << gs->optarg[g]);
            double c = cos(ang), s = sin(ang);
            /* Complex::operator*= (basic_host_types.hpp:69-73) */
            double a = st[2 * i11], b = st[2 * i11 + 1];
            st[2 * i11] = a * c - b * s;
            st[2 * i11 + 1] = a * s + b * c;
        }
        return 0;
    }
    case ORC_CNOT: { /* target q[0], control q[1]  (:46-64) */
        if (k != 2) return -1;
        for (uint64_t line = 0; line < cover; line++) {
            uint64_t i10 = insert_zeros(line, sorted, 2) | (1ull << lq[1]);
            uint64_t i11 = i10 | (1ull << lq[0]);
            double r = st[2 * i10], m = st[2 * i10 + 1];
            st[2 * i10] = st[2 * i11]; st[2 * i10 + 1] = st[2 * i11 + 1];
            st[2 * i11] = r; st[2 * i11 + 1] = m;
        }
        return 0;
    }
    case ORC_H: { /* a0' = (a0+a1)/sqrt2 ; a1' = a0' - a1*sqrt2  (:65-82) */
        if (k != 1) return -1;
        for (uint64_t line = 0; line < cover; line++) {
            uint64_t i0 = insert_zeros(line, sorted, 1);
            uint64_t i1 = i0 | (1ull << lq[0]);
            double r0 = (st[2 * i0] + st[2 * i1]) * ORC_SQRT2INV;
            double m0 = (st[2 * i0 + 1] + st[2 * i1 + 1]) * ORC_SQRT2INV;
            double r1 = r0 - st[2 * i1] * ORC_SQRT2;
            double m1 = m0 - st[2 * i1 + 1] * ORC_SQRT2;
            st[2 * i0] = r0; st[2 * i0 + 1] = m0;
            st[2 * i1] = r1; st[2 * i1 + 1] = m1;
        }
        return 0;
    }
    case ORC_DENSE: { /* sequential row sums in FP64  (:83-127) */
        if (gs->moff[g] < 0) return -1;
        const double* M = gs->mat + gs->moff[g];
        const uint64_t dim = 1ull << k;
        double* cache = (double*)malloc(sizeof(double) * 2 * dim);
        uint64_t* idx = (uint64_t*)malloc(sizeof(uint64_t) * dim);
        if (!cache || !idx) { free(cache); free(idx); return -2; }
        for (uint64_t line = 0; line < cover; line++) {
            uint64_t base = insert_zeros(line, sorted, k);
            for (uint64_t s = 0; s < dim; s++) {
                uint64_t li = base;
                for (int i = 0; i < k; i++) li += ((s >> i) & 1ull) << lq[i];
                idx[s] = li;
                cache[2 * s] = st[2 * li];
                cache[2 * s + 1] = st[2 * li + 1];
            }
            for (uint64_t r = 0; r < dim; r++) {
                double sr = 0.0, si = 0.0;
                for (uint64_t c = 0; c < dim; c++) {
                    double ma = M[2 * (r * dim + c)], mb = M[2 * (r * dim + c) + 1];
                    double va = cache[2 * c], vb = cache[2 * c + 1];
                    sr += ma * va - mb * vb;
                    si += ma * vb + mb * va;
                }
                st[2 * idx[r]] = sr;
                st[2 * idx[r] + 1] = si;
            }
        }
        free(cache);
        free(idx);
        return 0;
    }
    default:
        return -1; /* ids 6/7 (RX/RZ) are not computed by Kate either */
    }
}

/* CPUmergeGate, CPUcode/GateMerger.hpp:8-55: gates [first, last) composed into
 * one dense gate on the ascending union of their qubits.  Returns k. */
int orc_merge(const orc_gates* gs, int first, int last, int32_t* out_qubits, double* out_mat)
{
    uint64_t cover = 0;
    int maxq = 0;
    for (int g = first; g < last; g++)
        for (int i = 0; i < gs->nq[g]; i++) {
            int qq = gs->qubits[gs->qoff[g] + i];
            cover |= 1ull << qq;
            if (qq > maxq) maxq = qq;
        }
    int k = 0;
    int32_t* b2l = (int32_t*)calloc((size_t)maxq + 1, sizeof(int32_t));
    for (int qq = 0; qq <= maxq; qq++)
        if (cover >> qq & 1ull) { out_qubits[k] = qq; b2l[qq] = k; k++; }
    const uint64_t dim = 1ull << k;
    double* rows = (double*)calloc(2 * dim * dim, sizeof(double));
    for (uint64_t i = 0; i < dim; i++) rows[2 * (i * dim + i)] = 1.0;
    for (uint64_t i = 0; i < dim; i++)
        for (int g = first; g < last; g++)
            orc_apply_gate(gs, g, k, rows + 2 * i * dim, b2l);
    for (uint64_t r = 0; r < dim; r++) /* transpose: row i holds column i */
        for (uint64_t c = 0; c < dim; c++) {
            out_mat[2 * (r * dim + c)] = rows[2 * (c * dim + r)];
            out_mat[2 * (r * dim + c) + 1] = rows[2 * (c * dim + r) + 1];
        }
    free(rows);
    free(b2l);
    return k;
}

void orc_init_basis(int nq, double* st)
{
    memset(st, 0, sizeof(double) * 2 * (1ull << nq));
    st[0] = 1.0;
}

/* CPUcode/simulator.hpp:139-162: product state from (theta, phi) pairs. */
void orc_init_product(int nq, const double* tp, const int32_t* perm, double* st)
{
    double* a = (double*)malloc(sizeof(double) * 4 * (size_t)nq);
    for (int i = 0; i < nq; i++) {
        double th = tp[2 * i], ph = tp[2 * i + 1];
        double s = sin(th * ORC_PI / 2);
        a[2 * perm[i]] = cos(th * ORC_PI / 2);
        a[2 * perm[i] + 1] = 0.0;
        a[2 * (perm[i] + nq)] = cos(ph) * s;
        a[2 * (perm[i] + nq) + 1] = sin(ph) * s;
    }
    for (uint64_t i = 0; i < (1ull << nq); i++) {
        double r = 1.0, m = 0.0;
        for (int j = 0; j < nq; j++) {
            const double* f = a + 2 * (((i >> j) & 1ull) * (uint64_t)nq + (uint64_t)j);
            double t = r;
            r = r * f[0] - m * f[1];
            m = t * f[1] + m * f[0];
        }
        st[2 * i] = r;
        st[2 * i + 1] = m;
    }
    free(a);
}

void orc_swap_bits(int nq, double* st, int p1, int p2)
{
    if (p1 == p2) return;
    if (p1 > p2) { int t = p1; p1 = p2; p2 = t; }
    int pos[2] = {p1, p2};
    for (uint64_t line = 0; line < (1ull << (nq - 2)); line++) {
        uint64_t b = insert_zeros(line, pos, 2);
        uint64_t x = b | (1ull << p1), y = b | (1ull << p2);
        double r = st[2 * x], m = st[2 * x + 1];
        st[2 * x] = st[2 * y]; st[2 * x + 1] = st[2 * y + 1];
        st[2 * y] = r; st[2 * y + 1] = m;
    }
}

/* ---- executeCommand (CPUcode/simulator.hpp:198-291) ---- */
typedef struct {
    const orc_gates* gs;
    double* state;
    int nq, gq, g0, g1;
    const int* gqubits;
    const int32_t* b2l;
    uint64_t lo, hi;
    int err;
} orc_work;

static void* orc_thread(void* arg)
{
    orc_work* w = (orc_work*)arg;
    const uint64_t dim = 1ull << w->gq;
    double* local = (double*)malloc(sizeof(double) * 2 * dim);
    if (!local) { w->err = -2; return NULL; }
    for (uint64_t grp = w->lo; grp < w->hi; grp++) {
        uint64_t base = insert_zeros(grp, w->gqubits, w->gq);
        for (uint64_t line = 0; line < dim; line++) {
            uint64_t at = base;
            for (int i = 0; i < w->gq; i++) at += ((line >> i) & 1ull) << w->gqubits[i];
            local[2 * line] = w->state[2 * at];
            local[2 * line + 1] = w->state[2 * at + 1];
        }
        for (int g = w->g0; g < w->g1; g++)
            if (orc_apply_gate(w->gs, g, w->gq, local, w->b2l) != 0) w->err = -1;
        for (uint64_t line = 0; line < dim; line++) {
            uint64_t at = base;
            for (int i = 0; i < w->gq; i++) at += ((line >> i) & 1ull) << w->gqubits[i];
            w->state[2 * at] = local[2 * line];
            w->state[2 * at + 1] = local[2 * line + 1];
        }
    }
    free(local);
    return NULL;
}

static int orc_group(const orc_program* p, int grp, double* st, int nthreads)
{
    const int nq = p->nqubits;
    uint64_t set = p->group_mask[grp];
    /* pad with the lowest free qubits up to 14 (simulator.hpp:201-205) */
    for (int l = 0; l < nq; l++) {
        int cnt = __builtin_popcountll(set);
        if (cnt >= 14 || cnt == nq) break;
        set |= 1ull << l;
    }
    int gqubits[64], gq = 0;
    int32_t b2l[64];
    memset(b2l, 0, sizeof(b2l));
    for (int q = 0; q < nq; q++)
        if (set >> q & 1ull) { b2l[q] = gq; gqubits[gq++] = q; }
    int g0 = grp == 0 ? 0 : p->group_end[grp - 1];
    int g1 = p->group_end[grp];
    for (int g = g0; g < g1; g++) /* every gate qubit must be in the group */
        for (int i = 0; i < p->gates.nq[g]; i++)
            if (!(set >> p->gates.qubits[p->gates.qoff[g] + i] & 1ull)) return -3;

    uint64_t ngroups = 1ull << (nq - gq);
    uint64_t nt = (uint64_t)(nthreads < 1 ? 1 : nthreads);
    if (nt > ngroups) nt = ngroups;
    /* ceil split: every group is computed whatever the thread count (the
     * reference's floor split drops the remainder, simulator.hpp:280-287) */
    uint64_t per = (ngroups + nt - 1) / nt;
    orc_work* w = (orc_work*)calloc(nt, sizeof(orc_work));
    pthread_t* th = (pthread_t*)calloc(nt, sizeof(pthread_t));
    for (uint64_t t = 0; t < nt; t++) {
        w[t].gs = &p->gates; w[t].state = st; w[t].nq = nq; w[t].gq = gq;
        w[t].g0 = g0; w[t].g1 = g1; w[t].gqubits = gqubits; w[t].b2l = b2l;
        w[t].lo = t * per;
        w[t].hi = (t + 1) * per < ngroups ? (t + 1) * per : ngroups;
        if (w[t].lo > w[t].hi) w[t].lo = w[t].hi;
        if (nt == 1) orc_thread(&w[t]);
        else pthread_create(&th[t], NULL, orc_thread, &w[t]);
    }
    int err = 0;
    for (uint64_t t = 0; t < nt; t++) {
        if (nt > 1) pthread_join(th[t], NULL);
        if (w[t].err) err = w[t].err;
    }
    free(w);
    free(th);
    return err;
}

/* CPUcode/simulator.hpp:71-99, with swap instructions performed. */
int orc_run(const orc_program* p, double* st, int nthreads)
{
    int grp = 0;
    for (int i = 0; i < p->ninstr; i++) {
        if (p->instr_kind[i] == 0) {
            const int32_t* pr = p->pairs + p->instr_poff[i];
            for (int j = 0; j < p->instr_npairs[i]; j++) orc_swap_bits(p->nqubits, st, pr[2 * j], pr[2 * j + 1]);
        } else {
            if (grp >= p->ngroups) return -4;
            int e = orc_group(p, grp, st, nthreads);
            if (e) return e;
            grp++;
        }
    }
    return 0;
}

/* CPUcode/simulator.hpp:163-197. */
void orc_measure(int nq, const double* st, const int32_t* finv, double* out)
{
    double* m = (double*)calloc(4 * (size_t)nq, sizeof(double));
    for (uint64_t i = 0; i < (1ull << nq); i++)
        for (int q = 0; q < nq; q++) {
            double* acc = m + 2 * (2 * q + ((i >> q) & 1ull));
            acc[0] += st[2 * i];
            acc[1] += st[2 * i + 1];
        }
    const double den = pow(ORC_SQRT2, (double)nq);
    for (int q = 0; q < nq; q++) {
        /* Complex::operator/ with a real divisor (basic_host_types.hpp:54-57) */
        double nn = den * den + 0.0 * 0.0;
        double a0 = (m[4 * q] * den + m[4 * q + 1] * 0.0) / nn;
        double b0 = (den * m[4 * q + 1] - m[4 * q] * 0.0) / nn;
        double a1 = (m[4 * q + 2] * den + m[4 * q + 3] * 0.0) / nn;
        double b1 = (den * m[4 * q + 3] - m[4 * q + 2] * 0.0) / nn;
        double n0 = sqrt(a0 * a0 + b0 * b0), n1 = sqrt(a1 * a1 + b1 * b1);
        double th, ph;
        if (n0 < 0.0000000000001) { th = 1; ph = 0; }
        else if (n1 < 0.00000000000001) { th = 0; ph = 0; }
        else {
            th = atan(n1 / n0) / (ORC_PI / 2);
            double d = a0 * a0 + b0 * b0;
            double qa = (a1 * a0 + b1 * b0) / d;
            double qb = (a0 * b1 - a1 * b0) / d;
            ph = atan(qb / qa); /* Complex::angle is atan, not atan2 (:31-33) */
        }
        out[2 * finv[q]] = th;
        out[2 * finv[q] + 1] = ph;
    }
    free(m);
}

void orc_expect_z(int nq, const double* st, int m, const uint64_t* masks, double* out)
{
    for (int j = 0; j < m; j++) out[j] = 0.0;
    for (uint64_t i = 0; i < (1ull << nq); i++) {
        double p = st[2 * i] * st[2 * i] + st[2 * i + 1] * st[2 * i + 1];
        for (int j = 0; j < m; j++) out[j] += (__builtin_popcountll(i & masks[j]) & 1) ? -p : p;
    }
}
